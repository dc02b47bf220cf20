// Per-frame CPU front-end: camera, portal visibility, per-polygon surface setup, draw ordering and
// draw setup. It emits the command buffers of include/sw_gpu.h for one view.
//
// Restates, in order of execution:
//   matrix.rs:13-109                               camera matrices
//   map_visibility_calculator.rs:52-326            portal flood with octagon bounds
//   partial_renderer.rs:1327-1638                  prepare_polygons_surfaces / prepare_polygon_surface
//   renderer_utils.rs:10-58, :60-242               projection + clipping, interpolation-mode tests, mip
//   partial_renderer.rs:1640-1820                  per-surface parameters of build_polygons_surfaces
//   partial_renderer.rs:3877-3985                  polygon_is_affected_by_light
//   partial_renderer.rs:2311-2448                  draw_skybox
//   partial_renderer.rs:2450-2576, :4011-4142      draw_tree / draw_leaf / draw_polygon
//   common/clipping.rs, common/clipping_polygon.rs 3D / 2D clipping, ClippingPolygon
#include <algorithm>
#include <cstring>
#include <new>
#include <random>

#include "host_internal.h"
#include "host_frame_state.h"

using namespace swh;

namespace swh
{


// ---------------------------------------------------------------------------------------------
// clipping.rs
// ---------------------------------------------------------------------------------------------
V3 line_z_intersection(V3 v0, V3 v1, float z)
{
	const float dist0 = v0.z - z, dist1 = v1.z - z, dist_sum = v1.z - v0.z;
	const float k0 = dist0 / dist_sum, k1 = dist1 / dist_sum;
	return v0 * k1 - v1 * k0;
}

uint32_t clip_3d_polygon_by_z_plane(const V3* poly, uint32_t n, float z_dist, V3* out, uint32_t cap)
{
	if(n == 0)
		return 0;
	const V3* prev = &poly[n - 1];
	uint32_t count = 0;
	for(uint32_t i = 0; i < n; ++i)
	{
		const V3& v = poly[i];
		if(v.z > z_dist)
		{
			if(prev->z < z_dist)
			{
				out[count++] = line_z_intersection(*prev, v, z_dist);
				if(count == cap)
					break;
			}
			out[count++] = v;
			if(count == cap)
				break;
		}
		else if(v.z == z_dist)
		{
			out[count++] = v;
			if(count == cap)
				break;
		}
		else if(prev->z > z_dist)
		{
			out[count++] = line_z_intersection(*prev, v, z_dist);
			if(count == cap)
				break;
		}
		prev = &v;
	}
	return count;
}

V2 line_line_intersection(V2 v0, V2 v1, V3 line)
{
	const float dist0 = dot(v0, v2(line.x, line.y)) - line.z, dist1 = dot(v1, v2(line.x, line.y)) - line.z;
	const float dist_sum = dist1 - dist0;
	const float k0 = dist0 / dist_sum, k1 = dist1 / dist_sum;
	return v0 * k1 - v1 * k0;
}

uint32_t clip_2d_polygon(const V2* poly, uint32_t n, V3 line, V2* out, uint32_t cap)
{
	if(n == 0)
		return 0;
	const V2* prev = &poly[n - 1];
	float prev_dist = dot(*prev, v2(line.x, line.y)) - line.z;
	uint32_t count = 0;
	for(uint32_t i = 0; i < n; ++i)
	{
		const V2& v = poly[i];
		const float dist = dot(v, v2(line.x, line.y)) - line.z;
		if(dist > 0.0f)
		{
			if(prev_dist < 0.0f)
			{
				out[count++] = line_line_intersection(*prev, v, line);
				if(count == cap)
					break;
			}
			out[count++] = v;
			if(count == cap)
				break;
		}
		else if(dist == 0.0f)
		{
			out[count++] = v;
			if(count == cap)
				break;
		}
		else if(prev_dist > 0.0f)
		{
			out[count++] = line_line_intersection(*prev, v, line);
			if(count == cap)
				break;
		}
		prev = &v;
		prev_dist = dist;
	}
	return count;
}

// renderer_utils.rs:10-58
uint32_t project_and_clip_polygon(const V3 clip_planes[8], const V3* verts, uint32_t n, V2* out)
{
	uint32_t count = std::min(n, MAX_VERTICES);
	V3 zc[MAX_VERTICES];
	count = clip_3d_polygon_by_z_plane(verts, count, Z_NEAR, zc, MAX_VERTICES);
	if(count < 3)
		return count;
	for(uint32_t i = 0; i < count; ++i)
		out[i] = v2(zc[i].x, zc[i].y) / zc[i].z;
	V2 tmp[MAX_VERTICES];
	for(int i = 0; i < 4; ++i)
	{
		count = clip_2d_polygon(out, count, clip_planes[i * 2], tmp, MAX_VERTICES);
		if(count < 3)
			return count;
		count = clip_2d_polygon(tmp, count, clip_planes[i * 2 + 1], out, MAX_VERTICES);
		if(count < 3)
			return count;
	}
	return count;
}

// ---------------------------------------------------------------------------------------------
// equations.rs
// ---------------------------------------------------------------------------------------------
DepthEq depth_from_plane(V4 p)
{
	const float w = -p.w;
	return DepthEq{p.x / w, p.y / w, p.z / w};
}

TcEq tc_from_depth_and_planes(const DepthEq& d, const V4 t[2])
{
	TcEq e;
	for(int i = 0; i < 2; ++i)
	{
		e.dx[i] = t[i].x + t[i].w * d.dx;
		e.dy[i] = t[i].y + t[i].w * d.dy;
		e.k[i] = t[i].z + t[i].w * d.k;
	}
	return e;
}

DepthEq scale(const DepthEq& d, float s) { return DepthEq{d.dx * s, d.dy * s, d.k * s}; }
TcEq scale(const TcEq& e, float s)
{
	TcEq r;
	for(int i = 0; i < 2; ++i)
	{
		r.dx[i] = e.dx[i] * s;
		r.dy[i] = e.dy[i] * s;
		r.k[i] = e.k[i] * s;
	}
	return r;
}

// ---------------------------------------------------------------------------------------------
// renderer_utils.rs:60-242
// ---------------------------------------------------------------------------------------------
const float TC_ERROR_THRESHOLD = 0.75f;

bool affine_may_be_used(const DepthEq& de, const TcEq& te, V2 min_p, V2 max_p)
{
	const V2 edge = max_p - min_p;
	const float edge_square_len = dot(edge, edge);
	if(edge_square_len == 0.0f)
		return true;
	const float edge_len = std::sqrt(edge_square_len);
	const V2 en = edge / edge_len;
	const float inv_z_clamp = 1.0f / float(1 << 20);
	const float min_point_inv_z = std::fmax(de.sample(min_p), inv_z_clamp);
	const float max_point_inv_z = std::fmax(de.sample(max_p), inv_z_clamp);
	(void)max_point_inv_z;
	const float a = dot(v2(de.dx, de.dy), en);
	const float b = min_point_inv_z;
	if(std::fabs(a) < 1.0e-10f)
		return true;
	const float b_div_a = b / a;
	const float max_diff_point = std::sqrt((0.0f + b_div_a) * (edge_len + b_div_a)) - b_div_a;
	const float max_diff_point_inv_z = a * max_diff_point + b;
	for(int i = 0; i < 2; ++i)
	{
		const float min_point_tc = te.dx[i] * min_p.x + te.dy[i] * min_p.y + te.k[i];
		const float max_point_tc = te.dx[i] * max_p.x + te.dy[i] * max_p.y + te.k[i];
		const float tc_a = dot(v2(te.dx[i], te.dy[i]), en);
		const float tc_b = min_point_tc;
		const float min_mul = min_point_tc / min_point_inv_z;
		const float max_mul = max_point_tc / std::fmax(de.sample(max_p), inv_z_clamp);
		const float real = (tc_a * max_diff_point + tc_b) / max_diff_point_inv_z;
		const float approx = min_mul + (max_mul - min_mul) * (max_diff_point - 0.0f) / (edge_len - 0.0f);
		if(std::fabs(real - approx) > TC_ERROR_THRESHOLD)
			return false;
	}
	return true;
}

bool line_z_may_be_used(const DepthEq& de, const TcEq& te, V2 max_p, float min_x, float max_x)
{
	if(max_x - min_x < 1.0f)
		return true;
	const float a = de.dx;
	const float b = de.dy * max_p.y + de.k;
	if(std::fabs(a) < 1.0e-10f)
		return true;
	const float b_div_a = b / a;
	const float max_diff_x = std::sqrt((min_x + b_div_a) * (max_x + b_div_a)) - b_div_a;
	const float max_diff_point_inv_z = a * max_diff_x + b;
	const float inv_z_at_min_x = a * min_x + b;
	const float inv_z_at_max_x = a * max_x + b;
	const float almost_zero = 1e-20f;
	if(inv_z_at_min_x <= almost_zero || inv_z_at_max_x <= almost_zero || max_diff_point_inv_z <= almost_zero)
		return false;
	for(int i = 0; i < 2; ++i)
	{
		const float ta = te.dx[i];
		const float tb = te.dy[i] * max_p.y + te.k[i];
		const float tc_at_min_x = (ta * min_x + tb) / inv_z_at_min_x;
		const float tc_at_max_x = (ta * max_x + tb) / inv_z_at_max_x;
		const float real = (ta * max_diff_x + tb) / max_diff_point_inv_z;
		const float approx = tc_at_min_x + (tc_at_max_x - tc_at_min_x) * (max_diff_x - min_x) / (max_x - min_x);
		if(std::fabs(real - approx) > TC_ERROR_THRESHOLD)
			return false;
	}
	return true;
}

uint32_t calculate_mip(const V2* points, uint32_t n, const DepthEq& de, const TcEq& te, float mip_bias)
{
	V2 mip_point = points[0];
	float mip_point_inv_z = 0.0f;
	for(uint32_t i = 0; i < n; ++i)
	{
		const float inv_z = de.sample(points[i]);
		if(inv_z > mip_point_inv_z)
		{
			mip_point_inv_z = inv_z;
			mip_point = points[i];
		}
	}
	const float z_2 = 1.0f / (mip_point_inv_z * mip_point_inv_z);
	const float z_4 = z_2 * z_2;
	float d_tc_2[2];
	for(int i = 0; i < 2; ++i)
	{
		const float d_tc_dx = te.dx[i] * (de.k + de.dy * mip_point.y) - (te.k[i] + te.dy[i] * mip_point.y) * de.dx;
		const float d_tc_dy = te.dy[i] * (de.k + de.dx * mip_point.x) - (te.k[i] + te.dx[i] * mip_point.x) * de.dy;
		d_tc_2[i] = (d_tc_dx * d_tc_dx + d_tc_dy * d_tc_dy) * z_4;
	}
	const float max_d_tc_2 = std::fmax(d_tc_2[0], d_tc_2[1]);
	const float mip_f = std::log2(max_d_tc_2) * 0.5f + mip_bias;
	const float c = std::ceil(mip_f);
	// Rust `as i32`: saturating, NaN -> 0
	int32_t mi = std::isnan(c) ? 0 : (c >= 2147483647.0f ? INT32_MAX : (c <= -2147483648.0f ? INT32_MIN : int32_t(c)));
	return uint32_t(std::max(0, std::min(mi, int32_t(MAX_MIP))));
}

// f32 -> Fixed16 with x86 cvttss2si semantics (fixed_math.rs:36-39)
int32_t f32_to_fixed16_unchecked(float x)
{
	const float s = x * 65536.0f;
	if(!(s >= -2147483648.0f && s < 2147483648.0f))
		return INT32_MIN;
	return int32_t(s);
}

int32_t saturating_f32_to_i32(float x)
{
	if(std::isnan(x))
		return 0;
	if(x >= 2147483647.0f)
		return INT32_MAX;
	if(x <= -2147483648.0f)
		return INT32_MIN;
	return int32_t(x);
}

V3 view_transform(const M4& m, V3 v) { return truncate(mul(m, extend(v, 1.0f))); }

// Quaternion (cgmath) just far enough for the camera and the skybox
struct Quat
{
	float s;
	V3 v;
};
Quat quat_mul(Quat a, Quat b)
{
	return Quat{a.s * b.s - a.v.x * b.v.x - a.v.y * b.v.y - a.v.z * b.v.z,
		v3(a.s * b.v.x + a.v.x * b.s + a.v.y * b.v.z - a.v.z * b.v.y, a.s * b.v.y + a.v.y * b.s + a.v.z * b.v.x - a.v.x * b.v.z,
			a.s * b.v.z + a.v.z * b.s + a.v.x * b.v.y - a.v.y * b.v.x)};
}
Quat from_angle_z(float a) { return Quat{std::cos(a * 0.5f), v3(0.0f, 0.0f, std::sin(a * 0.5f))}; }
Quat from_angle_y(float a) { return Quat{std::cos(a * 0.5f), v3(0.0f, std::sin(a * 0.5f), 0.0f)}; }
M4 quat_to_m4(Quat q)
{
	const float x2 = q.v.x + q.v.x, y2 = q.v.y + q.v.y, z2 = q.v.z + q.v.z;
	const float xx2 = x2 * q.v.x, xy2 = x2 * q.v.y, xz2 = x2 * q.v.z, yy2 = y2 * q.v.y, yz2 = y2 * q.v.z, zz2 = z2 * q.v.z;
	const float sy2 = y2 * q.s, sz2 = z2 * q.s, sx2 = x2 * q.s;
	M4 m = identity();
	m.c[0] = v4(1.0f - yy2 - zz2, xy2 + sz2, xz2 - sy2, 0.0f);
	m.c[1] = v4(xy2 - sz2, 1.0f - xx2 - zz2, yz2 + sx2, 0.0f);
	m.c[2] = v4(xz2 + sy2, yz2 - sx2, 1.0f - xx2 - yy2, 0.0f);
	return m;
}

// matrix.rs:82-109 (complete_view_matrix)
Camera complete_view(V3 position, const M4& rotation_matrix, V2 fov_scale, float width, float height)
{
	const M4 translate = from_translation(v3(-position.x, -position.y, -position.z));
	const M4 perspective = from_nonuniform_scale(fov_scale.x, fov_scale.y, 1.0f);
	const M4 resize = from_nonuniform_scale(width * 0.5f, height * 0.5f, 1.0f);
	M4 shift = identity();
	shift.c[2].x = width * 0.5f;
	shift.c[2].y = height * 0.5f;
	Camera c;
	c.position = position;
	c.view_matrix = mul(mul(mul(mul(shift, resize), perspective), rotation_matrix), translate);
	if(!invert(transpose(c.view_matrix), c.planes_matrix))
		c.planes_matrix = identity();
	return c;
}

// matrix.rs:13-80
Camera build_camera(V3 position, float azimuth, float elevation, float fov, float width, float height)
{
	const Quat rotation = quat_mul(from_angle_z(azimuth + 3.14159265f * 0.5f), from_angle_y(-elevation));
	const float inv_half_fov_tan = 1.0f / std::tan(fov * 0.5f);
	const float aspect = width / height;
	const V2 fov_scale = v2(inv_half_fov_tan / aspect, inv_half_fov_tan);
	M4 basis_change = identity();
	basis_change.c[0].x = 0.0f;
	basis_change.c[1].y = 0.0f;
	basis_change.c[2].z = 0.0f;
	basis_change.c[0].z = 1.0f;
	basis_change.c[1].x = -1.0f;
	basis_change.c[2].y = -1.0f;
	const Quat conj = Quat{rotation.s, v3(-rotation.v.x, -rotation.v.y, -rotation.v.z)};
	const M4 rotation_matrix = mul(basis_change, quat_to_m4(conj));
	return complete_view(position, rotation_matrix, fov_scale, width, height);
}

// light.rs:88-128 (get_cube_map_side_matrix): the side's rotation columns; negated unit vectors
// keep their negative zeros
M4 cube_side_rotation(uint32_t side)
{
	const V4 px = v4(1.0f, 0.0f, 0.0f, 0.0f), nx = v4(-1.0f, -0.0f, -0.0f, -0.0f);
	const V4 py = v4(0.0f, 1.0f, 0.0f, 0.0f), ny = v4(-0.0f, -1.0f, -0.0f, -0.0f);
	const V4 pz = v4(0.0f, 0.0f, 1.0f, 0.0f), nz = v4(-0.0f, -0.0f, -1.0f, -0.0f);
	static const int cols[6][3] = {{3, -1, -2}, {-3, 1, -2}, {1, 3, -2}, {-1, -3, -2}, {1, 2, 3}, {1, -2, -3}};
	const V4 pos[3] = {px, py, pz}, neg[3] = {nx, ny, nz};
	M4 m = identity();
	for(int c = 0; c < 3; ++c)
	{
		const int k = cols[side % 6][c]; // +-(axis + 1)
		m.c[c] = k > 0 ? pos[k - 1] : neg[-k - 1];
	}
	return m;
}

} // namespace swh

namespace
{

// PolygonBasisVecs::form_plane_and_tex_coord_equation (surfaces.rs:18-47)
swh_frame::Basis form_basis(const Plane& plane, const Plane tc[2])
{
	M4 b;
	b.c[0] = extend(tc[0].vec, tc[0].dist);
	b.c[1] = extend(tc[1].vec, tc[1].dist);
	b.c[2] = extend(plane.vec, -plane.dist);
	b.c[3] = v4(0, 0, 0, 1);
	M4 inv;
	if(!invert(transpose(b), inv))
		return swh_frame::Basis{v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, 0), v3(0, 0, 1)};
	swh_frame::Basis out;
	out.u = truncate(inv.c[0]);
	out.v = truncate(inv.c[1]);
	out.start = truncate(inv.c[3]);
	out.normal = plane.vec / std::fmax(length(plane.vec), 1.0f / float(1 << 30));
	return out;
}

void ensure_basis(swh_frame& f, const swh_map& m)
{
	if(f.basis_map == &m && f.basis.size() == m.polygons.size())
		return;
	f.basis_map = &m;
	f.basis.resize(m.polygons.size());
	f.poly.assign(m.polygons.size(), swh_frame::PolyData{});
	f.vertices_transformed.assign(m.vertices.size(), v3(0, 0, 0));
	f.leaf_visible_frame.assign(m.leafs.size(), 0);
	f.leaf_bounds.assign(m.leafs.size(), ClippingPolygon{});
	f.portal_visible_frame.assign(m.portals.size(), 0);
	f.portal_has_projection.assign(m.portals.size(), 0);
	f.portal_projection.assign(m.portals.size(), ClippingPolygon{});
	f.frame_number = 0;
	for(size_t i = 0; i < m.polygons.size(); ++i)
	{
		const Polygon& p = m.polygons[i];
		f.basis[i] = form_basis(p.plane, p.tex_coord_equation);
	}
}

// map_visibility_calculator.rs:286-326
bool project_portal(const swh_map& m, const Portal& portal, const M4& view, ClippingPolygon& out)
{
	uint32_t n = std::min(portal.num_vertices, MAX_VERTICES);
	V3 t[MAX_VERTICES], zc[MAX_VERTICES];
	for(uint32_t i = 0; i < n; ++i)
		t[i] = view_transform(view, m.vertices[portal.first_vertex + i]);
	n = clip_3d_polygon_by_z_plane(t, n, 1.0f / 4096.0f, zc, MAX_VERTICES);
	if(n < 3)
		return false;
	out = ClippingPolygon::from_point(v2(zc[0].x, zc[0].y) / zc[0].z);
	for(uint32_t i = 1; i < n; ++i)
		out.extend_point(v2(zc[i].x, zc[i].y) / zc[i].z);
	return true;
}

// mark_reachable_leafs_iterative (map_visibility_calculator.rs:156-267): breadth-first over the
// leaf graph from the start leaves (each with the frame bounds), through portals facing the camera,
// each leaf's bounds the union of the projected portals' bounds clipped by the leaf it is seen from.
void mark_reachable_leafs(swh_frame& f, const swh_map& m, const Camera& cam, const ClippingPolygon& frame_bounds,
	const uint32_t* start_leafs, uint32_t num_start)
{
	const uint32_t frame = f.frame_number;
	std::vector<uint32_t> cur, next;
	for(uint32_t i = 0; i < num_start; ++i)
	{
		cur.push_back(start_leafs[i]);
		f.leaf_bounds[start_leafs[i]] = frame_bounds;
		f.leaf_visible_frame[start_leafs[i]] = frame;
	}
	int depth = 0;
	while(!cur.empty())
	{
		for(uint32_t leaf : cur)
		{
			const ClippingPolygon leaf_bounds = f.leaf_bounds[leaf];
			const Leaf& lv = m.leafs[leaf];
			for(uint32_t k = 0; k < lv.num_leaf_portals; ++k)
			{
				const uint32_t pi = m.leafs_portals[lv.first_leaf_portal + k];
				const Portal& portal = m.portals[pi];
				const float portal_plane_pos = mul(cam.planes_matrix, extend(portal.plane.vec, -portal.plane.dist)).w;
				uint32_t next_leaf;
				if(portal.leafs[0] == leaf)
				{
					if(portal_plane_pos <= 0.0f)
						continue;
					next_leaf = portal.leafs[1];
				}
				else
				{
					if(portal_plane_pos >= 0.0f)
						continue;
					next_leaf = portal.leafs[0];
				}
				if(f.portal_visible_frame[pi] != frame)
				{
					f.portal_visible_frame[pi] = frame;
					f.portal_has_projection[pi] = project_portal(m, portal, cam.view_matrix, f.portal_projection[pi]) ? 1 : 0;
				}
				if(!f.portal_has_projection[pi])
					continue;
				ClippingPolygon bi = f.portal_projection[pi];
				bi.intersect(leaf_bounds);
				if(!bi.valid())
					continue;
				if(f.leaf_visible_frame[next_leaf] != frame)
				{
					f.leaf_visible_frame[next_leaf] = frame;
					f.leaf_bounds[next_leaf] = bi;
				}
				else
				{
					if(f.leaf_bounds[next_leaf].contains(bi))
						continue;
					f.leaf_bounds[next_leaf].extend(bi);
				}
				next.push_back(next_leaf);
			}
		}
		cur.clear();
		std::swap(cur, next);
		if(++depth > 1024)
			break;
	}
}

// update_visibility (map_visibility_calculator.rs:52-98): the camera's leaf and the leaves behind portals
// the camera (almost) lies on
void update_visibility(swh_frame& f, const swh_map& m, const Camera& cam, const ClippingPolygon& frame_bounds)
{
	const uint32_t current_leaf = find_leaf(m, cam.position);
	uint32_t start_leafs[32];
	uint32_t num_start = 1;
	start_leafs[0] = current_leaf;
	const Leaf& leaf_value = m.leafs[current_leaf];
	for(uint32_t k = 0; k < leaf_value.num_leaf_portals; ++k)
	{
		const Portal& p = m.portals[m.leafs_portals[leaf_value.first_leaf_portal + k]];
		const float scaled_dist = dot(p.plane.vec, cam.position) - p.plane.dist;
		const float eps = Z_NEAR * 2.0f;
		if(std::fabs(scaled_dist) <= eps * length(p.plane.vec))
		{
			start_leafs[num_start++] = p.leafs[0] == current_leaf ? p.leafs[1] : p.leafs[0];
			if(num_start == 32)
				break;
		}
	}
	mark_reachable_leafs(f, m, cam, frame_bounds, start_leafs, num_start);
	// is_inside_leaf_volume (:269-283)
	f.inside_leaf_volume = true;
	for(uint32_t k = 0; k < leaf_value.num_polygons; ++k)
	{
		const Polygon& p = m.polygons[leaf_value.first_polygon + k];
		if(mul(cam.planes_matrix, extend(p.plane.vec, -p.plane.dist)).w < 0.0f)
		{
			f.inside_leaf_volume = false;
			break;
		}
	}
}

// update_visibility_with_start_leafs (map_visibility_calculator.rs:100-114): portal and mirror views
// start from the leaves of the portal; whether the camera is inside a leaf's volume is not known
void update_visibility_with_start_leafs(swh_frame& f, const swh_map& m, const Camera& cam, const ClippingPolygon& frame_bounds,
	const std::vector<uint32_t>& start_leafs)
{
	mark_reachable_leafs(f, m, cam, frame_bounds, start_leafs.data(), uint32_t(start_leafs.size()));
	f.inside_leaf_volume = true;
}

// partial_renderer.rs:1464-1638
void prepare_polygon_surface(swh_frame& f, const swh_map& m, const Camera& cam, const V3 clip_planes[8], uint64_t& offset,
	uint32_t pi, float mip_bias)
{
	swh_frame::PolyData& pd = f.poly[pi];
	const Polygon& polygon = m.polygons[pi];
	const V4 plane_transformed = mul(cam.planes_matrix, extend(polygon.plane.vec, -polygon.plane.dist));
	if(plane_transformed.w <= 0.0f)
		return;
	const Material& material = m.materials[polygon.texture];
	for(uint32_t k = 0; k < polygon.num_vertices; ++k)
		f.vertices_transformed[polygon.first_vertex + k] = view_transform(cam.view_matrix, m.vertices[polygon.first_vertex + k]);
	V2 v2d[MAX_VERTICES];
	const uint32_t n = project_and_clip_polygon(clip_planes, &f.vertices_transformed[polygon.first_vertex], polygon.num_vertices, v2d);
	if(n < 3)
		return;
	if(material.sky)
	{
		ClippingPolygon b = ClippingPolygon::from_point(v2d[0]);
		for(uint32_t i = 1; i < n; ++i)
			b.extend_point(v2d[i]);
		if(f.have_sky)
			f.sky_bounds.extend(b);
		else
		{
			f.have_sky = true;
			f.sky_material = polygon.texture;
			f.sky_bounds = b;
		}
		return;
	}
	const DepthEq de = depth_from_plane(plane_transformed);
	const V4 tcb[2] = {mul(cam.planes_matrix, extend(polygon.tex_coord_equation[0].vec, polygon.tex_coord_equation[0].dist)),
		mul(cam.planes_matrix, extend(polygon.tex_coord_equation[1].vec, polygon.tex_coord_equation[1].dist))};
	const TcEq te = tc_from_depth_and_planes(de, tcb);
	const uint32_t mip = calculate_mip(v2d, n, de, te, mip_bias);
	const TcEq tes = scale(te, 1.0f / float(1 << mip));

	const float inf = float(1 << 29);
	const float max_z = float(1 << 16);
	float tc_min[2] = {inf, inf}, tc_max[2] = {-inf, -inf};
	for(uint32_t k = 0; k < n; ++k)
	{
		const float inv_z = std::fmax(de.sample(v2d[k]), 1.0f / max_z);
		const float z = 1.0f / inv_z;
		for(int i = 0; i < 2; ++i)
		{
			const float tc = z * (tes.dx[i] * v2d[k].x + tes.dy[i] * v2d[k].y + tes.k[i]);
			if(tc < tc_min[i])
				tc_min[i] = tc;
			if(tc > tc_max[i])
				tc_max[i] = tc;
		}
	}
	int32_t surface_tc_min[2], surface_size[2];
	for(int i = 0; i < 2; ++i)
	{
		const float eps = 1.0f / 32.0f;
		tc_min[i] += eps;
		tc_max[i] -= eps;
		const int32_t round_mask = ~(int32_t(LIGHTMAP_SCALE) - 1);
		const int32_t tc_min_round_down = (polygon.tex_coord_min[i] & round_mask) >> mip;
		const int32_t tc_max_round_up = ((polygon.tex_coord_max[i] + int32_t(LIGHTMAP_SCALE) - 1) & round_mask) >> mip;
		int32_t tc_min_int = std::max(saturating_f32_to_i32(std::floor(std::fmax(tc_min[i], -inf))), tc_min_round_down);
		int32_t tc_max_int = std::min(saturating_f32_to_i32(std::ceil(std::fmin(tc_max[i], inf))), tc_max_round_up);
		if(tc_min_int >= tc_max_int)
		{
			tc_min_int = std::min(tc_min_int, tc_max_round_up - 1);
			tc_max_int = tc_min_int + 1;
		}
		surface_tc_min[i] = tc_min_int;
		surface_size[i] = tc_max_int - tc_min_int;
	}
	pd.visible_frame = f.frame_number;
	pd.de = de;
	pd.te = tes;
	pd.surface_offset = offset;
	pd.size[0] = uint32_t(surface_size[0]);
	pd.size[1] = uint32_t(surface_size[1]);
	pd.mip = mip;
	pd.tc_min[0] = surface_tc_min[0];
	pd.tc_min[1] = surface_tc_min[1];
	offset += uint64_t(pd.size[0]) * pd.size[1];
	for(int i = 0; i < 2; ++i)
	{
		const float t = float(surface_tc_min[i]);
		pd.te.dx[i] -= t * de.dx;
		pd.te.dy[i] -= t * de.dy;
		pd.te.k[i] -= t * de.k;
	}
	f.visible_polygons.push_back(pi);
}

// partial_renderer.rs:3877-3985 (ShadowMap::None branch)
bool polygon_is_affected_by_light(const Polygon& polygon, const swh_frame::Basis& b, const sw_light& light)
{
	const V3 lp = v3(light.position[0], light.position[1], light.position[2]);
	const V3 d = lp - b.start;
	const float sd = dot(d, b.normal);
	if(sd <= 0.0f)
		return false;
	const float sq = light.radius * light.radius - sd * sd;
	if(sq <= 0.0f)
		return false;
	// invert the 3x3 (u, v, normal) and transpose
	M4 m3 = identity();
	m3.c[0] = extend(b.u, 0.0f);
	m3.c[1] = extend(b.v, 0.0f);
	m3.c[2] = extend(b.normal, 0.0f);
	M4 inv;
	if(!invert(m3, inv))
		return false;
	const M4 t = transpose(inv);
	const float tc0 = dot(d, truncate(t.c[0])), tc1 = dot(d, truncate(t.c[1]));
	const float r = std::sqrt(sq);
	const float u_r = r * (1.0f / std::sqrt(dot(b.u, b.u)));
	const float v_r = r * (1.0f / std::sqrt(dot(b.v, b.v)));
	return !(tc0 + u_r < float(polygon.tex_coord_min[0]) || tc1 + v_r < float(polygon.tex_coord_min[1]) ||
		tc0 - u_r > float(polygon.tex_coord_max[0]) || tc1 - v_r > float(polygon.tex_coord_max[1]));
}

// The frame's inline-model entities (frame_info.rs SubmodelEntity, positioned by the game in the
// reference): doors swing about the vertical axis through their centre, platforms rise and sink,
// both as functions of the frame time.
void submodel_entity(uint32_t k, float time, V3 centre, V3& position, Quat& rotation)
{
	if(k % 2 == 0)
	{
		const float a = 0.7f * std::sin(0.9f * time + 1.3f * float(k));
		rotation = Quat{std::cos(0.5f * a), v3(0.0f, 0.0f, std::sin(0.5f * a))};
		position = centre;
	}
	else
	{
		rotation = Quat{1.0f, v3(0.0f, 0.0f, 0.0f)};
		position = centre + v3(0.0f, 0.0f, 24.0f * std::sin(1.1f * time + 0.7f * float(k)));
	}
}

// The leaves a convex hull of points reaches: get_convex_hull_bsp_leafs (bsp_map_compact.rs:294-338),
// which InlineModelsIndex::position_model_r (inline_models_index.rs:126-157) applies to a transformed
// bbox's corners and DynamicObjectsIndex::position_portals (dynamic_objects_index.rs:144-157) to a
// portal's vertices.
void convex_hull_leafs_r(const swh_map& m, const V3* points, uint32_t n, uint32_t node_index, std::vector<uint32_t>& leafs)
{
	if(node_index >= FIRST_LEAF_INDEX)
	{
		leafs.push_back(node_index - FIRST_LEAF_INDEX);
		return;
	}
	const Node& node = m.nodes[node_index];
	uint32_t front = 0;
	for(uint32_t i = 0; i < n; ++i)
		if(dot(node.plane.vec, points[i]) > node.plane.dist)
			++front;
	if(front > 0)
		convex_hull_leafs_r(m, points, n, node.children[0], leafs);
	if(front < n)
		convex_hull_leafs_r(m, points, n, node.children[1], leafs);
}

void bbox_corners(const BBox& b, V3 out[8])
{
	// BBox::get_corners_vertices (bbox.rs:83-95)
	for(int i = 0; i < 8; ++i)
		out[i] = v3((i & 4) ? b.max.x : b.min.x, (i & 2) ? b.max.y : b.min.y, (i & 1) ? b.max.z : b.min.z);
}

// InlineModelsIndex::position_models + get_model_matrix (inline_models_index.rs:31-124, 181-196) and
// prepare_submodels (partial_renderer.rs:593-664): model matrices, the camera's matrices of each
// model, its leaves and the dynamic lights whose sphere's box touches its transformed bbox.
void prepare_submodels(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const Camera& cam)
{
	f.submodels.assign(m.submodels.size(), swh_frame::SubmodelFrame{});
	f.leaf_submodels.clear();
	if(m.submodels.empty())
		return;
	f.leaf_submodels.assign(m.leafs.size(), std::vector<uint32_t>());
	for(uint32_t k = 0; k < m.submodels.size(); ++k)
	{
		swh_frame::SubmodelFrame& s = f.submodels[k];
		const Submodel& sm = m.submodels[k];
		// get_submodel_bbox (bsp_map_compact.rs:236-255) extended by 1/4 (inline_models_index.rs:171-176)
		BBox b{v3(1e24f, 1e24f, 1e24f), v3(-1e24f, -1e24f, -1e24f)};
		for(uint32_t pi = sm.first_polygon; pi < sm.first_polygon + sm.num_polygons; ++pi)
		{
			const Polygon& p = m.polygons[pi];
			for(uint32_t v = 0; v < p.num_vertices; ++v)
			{
				const V3 q = m.vertices[p.first_vertex + v];
				b.min = v3(std::min(b.min.x, q.x), std::min(b.min.y, q.y), std::min(b.min.z, q.z));
				b.max = v3(std::max(b.max.x, q.x), std::max(b.max.y, q.y), std::max(b.max.z, q.z));
			}
		}
		const V3 eps = v3(0.25f, 0.25f, 0.25f);
		b.min = b.min - eps;
		b.max = b.max + eps;
		s.bbox = b;
		const V3 centre = (b.min + b.max) * 0.5f;
		V3 position;
		Quat rotation;
		submodel_entity(k, cfg.time, centre, position, rotation);
		s.model = mul(mul(from_translation(position), quat_to_m4(rotation)), from_translation(centre * -1.0f));
		if(!invert(transpose(s.model), s.world_planes))
			continue;
		s.view = mul(cam.view_matrix, s.model);
		s.planes = mul(cam.planes_matrix, s.world_planes);
		s.active = true;
		V3 corners[8];
		bbox_corners(b, corners);
		BBox world{v3(1e30f, 1e30f, 1e30f), v3(-1e30f, -1e30f, -1e30f)};
		for(V3& c : corners)
		{
			c = truncate(mul(s.model, extend(c, 1.0f)));
			world.min = v3(std::min(world.min.x, c.x), std::min(world.min.y, c.y), std::min(world.min.z, c.z));
			world.max = v3(std::max(world.max.x, c.x), std::max(world.max.y, c.y), std::max(world.max.z, c.z));
		}
		convex_hull_leafs_r(m, corners, 8, uint32_t(m.nodes.size() - 1), s.leafs);
		for(uint32_t l : s.leafs)
			f.leaf_submodels[l].push_back(k);
		for(uint32_t li = 0; li < f.lights.size(); ++li)
		{
			const sw_light& l = f.lights[li];
			if(l.position[0] - l.radius > world.max.x || l.position[0] + l.radius < world.min.x ||
				l.position[1] - l.radius > world.max.y || l.position[1] + l.radius < world.min.y ||
				l.position[2] - l.radius > world.max.z || l.position[2] + l.radius < world.min.z)
				continue;
			s.lights.push_back(li);
		}
	}
}

// The submodel part of prepare_polygons_surfaces (partial_renderer.rs:1362-1447): models in a visible
// leaf get their polygons' surfaces with the model's camera matrices, clipped by the union of the
// bounds of their visible leaves, and each visible polygon's basis from its transformed planes.
void prepare_submodel_surfaces(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const Camera& cam, uint64_t& offset)
{
	for(uint32_t k = 0; k < m.submodels.size(); ++k)
	{
		const swh_frame::SubmodelFrame& s = f.submodels[k];
		if(!s.active)
			continue;
		bool have = false;
		ClippingPolygon bounds{};
		for(uint32_t l : s.leafs)
		{
			if(f.leaf_visible_frame[l] != f.frame_number)
				continue;
			if(have)
				bounds.extend(f.leaf_bounds[l]);
			else
				bounds = f.leaf_bounds[l];
			have = true;
		}
		if(!have)
			continue;
		V3 cp[8];
		bounds.clip_planes(cp);
		Camera sc = cam;
		sc.view_matrix = s.view;
		sc.planes_matrix = s.planes;
		const Submodel& sm = m.submodels[k];
		for(uint32_t pi = sm.first_polygon; pi < sm.first_polygon + sm.num_polygons; ++pi)
		{
			prepare_polygon_surface(f, m, sc, cp, offset, pi, cfg.textures_mip_bias);
			if(f.poly[pi].visible_frame != f.frame_number)
				continue;
			const Polygon& p = m.polygons[pi];
			const V4 pt = mul(s.world_planes, extend(p.plane.vec, -p.plane.dist));
			const V4 t0 = mul(s.world_planes, extend(p.tex_coord_equation[0].vec, p.tex_coord_equation[0].dist));
			const V4 t1 = mul(s.world_planes, extend(p.tex_coord_equation[1].vec, p.tex_coord_equation[1].dist));
			const Plane tcs[2] = {Plane{truncate(t0), t0.w}, Plane{truncate(t1), t1.w}};
			f.basis[pi] = form_basis(Plane{truncate(pt), -pt.w}, tcs);
		}
	}
}

// partial_renderer.rs:1682-1799 (per-surface parameters)
void emit_surface(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, uint32_t pi)
{
	const Polygon& polygon = m.polygons[pi];
	const swh_frame::PolyData& pd = f.poly[pi];
	const Material& material = m.materials[polygon.texture];
	const swh_frame::Basis& basis = f.basis[pi];
	sw_surface_desc s;
	std::memset(&s, 0, sizeof(s));
	const float scale = float(1u << pd.mip);
	const V3 u = pd.mip ? basis.u * scale : basis.u, v = pd.mip ? basis.v * scale : basis.v;
	s.basis_u[0] = u.x, s.basis_u[1] = u.y, s.basis_u[2] = u.z;
	s.basis_v[0] = v.x, s.basis_v[1] = v.y, s.basis_v[2] = v.z;
	s.basis_start[0] = basis.start.x, s.basis_start[1] = basis.start.y, s.basis_start[2] = basis.start.z;
	s.basis_normal[0] = basis.normal.x, s.basis_normal[1] = basis.normal.y, s.basis_normal[2] = basis.normal.z;
	s.size[0] = pd.size[0];
	s.size[1] = pd.size[1];
	s.tc_start[0] = pd.tc_min[0];
	s.tc_start[1] = pd.tc_min[1];
	s.texture_level = material.first_level + pd.mip;
	for(int i = 0; i < 2; ++i)
	{
		const int32_t round_mask = ~(int32_t(LIGHTMAP_SCALE) - 1);
		s.lightmap_tc_shift[i] = uint32_t(pd.tc_min[i] - ((polygon.tex_coord_min[i] & round_mask) >> pd.mip));
	}
	s.lightmap_size[0] = uint32_t(((polygon.tex_coord_max[0] + int32_t(LIGHTMAP_SCALE) - 1) >> LIGHTMAP_SCALE_LOG2) -
		(polygon.tex_coord_min[0] >> LIGHTMAP_SCALE_LOG2) + 1);
	s.lightmap_size[1] = uint32_t(((polygon.tex_coord_max[1] + int32_t(LIGHTMAP_SCALE) - 1) >> LIGHTMAP_SCALE_LOG2) -
		(polygon.tex_coord_min[1] >> LIGHTMAP_SCALE_LOG2) + 1);
	s.lightmap_scale_log2 = LIGHTMAP_SCALE_LOG2 - pd.mip;
	s.lightmap_offset = polygon.lightmap_data_offset;
	s.flags = SW_SURF_DIFFUSE_LAYER;
	if(cfg.use_directional_lightmaps && !m.lm_dir.empty())
		s.flags |= SW_SURF_DIRECTIONAL;
	s.emissive_level = SW_NO_TEXTURE;
	s.pixels_offset = pd.surface_offset;
	// dynamic lights of the polygon's leaf, or of its submodel (:1706-1719), at most 6
	s.first_light = uint32_t(f.light_indices.size());
	if(!f.leaf_lights.empty())
	{
		const std::vector<uint32_t>* list = nullptr;
		if(m.polygon_leaf[pi] != NO_POLYGON_LEAF)
			list = &f.leaf_lights[m.polygon_leaf[pi]];
		else
			for(size_t k = 0; k < m.submodels.size() && k < f.submodels.size(); ++k)
				if(pi >= m.submodels[k].first_polygon && pi - m.submodels[k].first_polygon < m.submodels[k].num_polygons)
					list = &f.submodels[k].lights;
		static const std::vector<uint32_t> none;
		for(uint32_t li : list ? *list : none)
		{
			if(polygon_is_affected_by_light(polygon, basis, f.lights[li]))
			{
				f.light_indices.push_back(li);
				if(++s.num_lights == SW_MAX_POLYGON_LIGHTS)
					break;
			}
		}
	}
	f.surfaces.push_back(s);
}

} // namespace

namespace swh
{

// partial_renderer.rs:4011-4142
void draw_polygon(swh_frame& f, const V3 clip_planes[8], const V3* verts, uint32_t nverts, const DepthEq& de, const TcEq& te,
	const uint32_t tsize[2], uint64_t texture_offset, uint32_t texture_source, uint32_t blend, const int32_t clip_rect[4])
{
	if(nverts < 3)
		return;
	V2 v2d[MAX_VERTICES];
	const uint32_t n = project_and_clip_polygon(clip_planes, verts, nverts, v2d);
	if(n < 3)
		return;
	float min_inv_z = 1e24f, max_inv_z = -1e24f, min_x = 1e24f, max_x = -1e24f;
	V2 min_p = v2d[0], max_p = v2d[0];
	for(uint32_t i = 0; i < n; ++i)
	{
		const float inv_z = de.sample(v2d[i]);
		if(inv_z < min_inv_z)
		{
			min_inv_z = inv_z;
			min_p = v2d[i];
		}
		if(inv_z > max_inv_z)
		{
			max_inv_z = inv_z;
			max_p = v2d[i];
		}
		if(v2d[i].x < min_x)
			min_x = v2d[i].x;
		if(v2d[i].x > max_x)
			max_x = v2d[i].x;
	}
	sw_draw d;
	std::memset(&d, 0, sizeof(d));
	d.kind = SW_DRAW_POLYGON;
	d.blend = blend;
	d.texture_source = texture_source;
	d.first_vertex_int = uint32_t(f.vdata.size());
	d.num_vertices = n;
	for(uint32_t i = 0; i < n; ++i)
	{
		f.vdata.push_back(f32_to_fixed16_unchecked(v2d[i].x));
		f.vdata.push_back(f32_to_fixed16_unchecked(v2d[i].y));
	}
	d.texture_offset = texture_offset;
	d.texture_size[0] = int32_t(tsize[0]);
	d.texture_size[1] = int32_t(tsize[1]);
	for(int i = 0; i < 4; ++i)
		d.clip_rect[i] = clip_rect[i];
	DepthEq des = de;
	TcEq tes = te;
	if(affine_may_be_used(de, te, min_p, max_p))
		d.mode = SW_TC_AFFINE;
	else
	{
		const float z_scale = std::exp2(-5.0f - std::ceil(std::log2(std::fmax(max_inv_z, 1.0f / float(1 << 20)))));
		des = scale(de, z_scale);
		tes = scale(te, z_scale);
		d.mode = line_z_may_be_used(de, te, max_p, min_x, max_x) ? SW_TC_LINE_Z : SW_TC_PERSPECTIVE;
	}
	d.depth_equation[0] = des.dx;
	d.depth_equation[1] = des.dy;
	d.depth_equation[2] = des.k;
	for(int i = 0; i < 2; ++i)
	{
		d.tc_equation[i] = tes.dx[i];
		d.tc_equation[2 + i] = tes.dy[i];
		d.tc_equation[4 + i] = tes.k[i];
	}
	f.draws.push_back(d);
}

} // namespace swh

namespace
{

// partial_renderer.rs:2311-2448
void draw_skybox(swh_frame& f, const swh_map& m, const Camera& cam, const ClippingPolygon& viewport, float mip_bias,
	const int32_t clip_rect[4])
{
	if(!f.have_sky || m.sky_side_size == 0)
		return;
	ClippingPolygon bounds = f.sky_bounds;
	bounds.intersect(viewport);
	if(!bounds.valid())
		return;
	static const float BOX[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}, {1, -1, 1}, {1, 1, -1}, {1, 1, 1}};
	struct Side
	{
		int idx[4];
		V3 n, u, v;
	};
	const Side sides[6] = {{{0, 1, 3, 2}, v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, -1)}, {{4, 6, 7, 5}, v3(-1, 0, 0), v3(0, -1, 0), v3(0, 0, -1)},
		{{0, 4, 5, 1}, v3(0, 1, 0), v3(-1, 0, 0), v3(0, 0, -1)}, {{2, 3, 7, 6}, v3(0, -1, 0), v3(1, 0, 0), v3(0, 0, -1)},
		{{0, 2, 6, 4}, v3(0, 0, 1), v3(1, 0, 0), v3(0, -1, 0)}, {{1, 5, 7, 3}, v3(0, 0, -1), v3(1, 0, 0), v3(0, 1, 0)}};
	const float side_plane_dist = -1.0f, side_tc_shift = 1.0f;
	const M4 skybox_matrix = from_translation(cam.position); // identity rotation
	M4 skybox_matrix_inverse;
	if(!invert(transpose(skybox_matrix), skybox_matrix_inverse))
		return;
	const M4 sview = mul(cam.view_matrix, skybox_matrix);
	const M4 splanes = mul(cam.planes_matrix, skybox_matrix_inverse);
	const float SKY_SCALE = 16384.0f;
	V3 bt[8];
	for(int i = 0; i < 8; ++i)
		bt[i] = view_transform(sview, v3(BOX[i][0], BOX[i][1], BOX[i][2]) * SKY_SCALE);
	V3 clip_planes[8];
	bounds.clip_planes(clip_planes);
	uint64_t side_base = 0;
	uint64_t side_texels = 0;
	for(uint32_t mip = 0, s = m.sky_side_size; mip <= MAX_MIP; ++mip, s >>= 1)
		side_texels += uint64_t(s) * s;
	for(int side = 0; side < 6; ++side, side_base += side_texels)
	{
		const V3 vt[4] = {bt[sides[side].idx[0]], bt[sides[side].idx[1]], bt[sides[side].idx[2]], bt[sides[side].idx[3]]};
		const DepthEq de = depth_from_plane(mul(splanes, extend(sides[side].n, -side_plane_dist)));
		const float tc_scale = float(m.sky_side_size) * 0.5f;
		const V4 u4 = extend(sides[side].u, side_tc_shift), v4_ = extend(sides[side].v, side_tc_shift);
		const V4 tcb[2] = {mul(splanes, v4(u4.x * tc_scale, u4.y * tc_scale, u4.z * tc_scale, u4.w * tc_scale)),
			mul(splanes, v4(v4_.x * tc_scale, v4_.y * tc_scale, v4_.z * tc_scale, v4_.w * tc_scale))};
		const TcEq te = tc_from_depth_and_planes(de, tcb);
		V2 v2d[MAX_VERTICES];
		const uint32_t n = project_and_clip_polygon(clip_planes, vt, 4, v2d);
		if(n < 3)
			continue;
		const uint32_t mip = calculate_mip(v2d, n, de, te, mip_bias);
		const TcEq tes = scale(te, 1.0f / float(1 << mip));
		uint64_t off = side_base;
		uint32_t size = m.sky_side_size;
		for(uint32_t k = 0; k < mip; ++k)
		{
			off += uint64_t(size) * size;
			size >>= 1;
		}
		const uint32_t ts[2] = {size, size};
		draw_polygon(f, clip_planes, vt, 4, de, tes, ts, off, SW_TEXSRC_SKYBOX, SW_BLEND_NONE, clip_rect);
	}
}

} // namespace

// =============================================================================================
// C API
// =============================================================================================
namespace
{

struct ScreenRect
{
	float min_x, min_y, max_x, max_y;
};

// split_rect (rect_splitting.rs:23-78): halves along the longer side, the split at
// max * ratio0 + min * ratio1 (f32), the first half getting num_parts / 2 of the parts.
void split_rect(const ScreenRect& r, uint32_t num_parts, std::vector<ScreenRect>& out)
{
	if(num_parts <= 1)
	{
		if(out.size() < 64)
			out.push_back(r);
		return;
	}
	const uint32_t n0 = num_parts / 2, n1 = num_parts - n0;
	const float ratio0 = float(n0) / float(num_parts), ratio1 = float(n1) / float(num_parts);
	ScreenRect r0 = r, r1 = r;
	if(r.max_x - r.min_x >= r.max_y - r.min_y)
	{
		const float middle = r.max_x * ratio0 + r.min_x * ratio1;
		r0.max_x = middle;
		r1.min_x = middle;
	}
	else
	{
		const float middle = r.max_y * ratio0 + r.min_y * ratio1;
		r0.max_y = middle;
		r1.min_y = middle;
	}
	split_rect(r0, n0, out);
	split_rect(r1, n1, out);
}

// perform_rasterization_for_viewport_part (partial_renderer.rs:1285-1325) for one viewport part:
// the skybox, then the BSP tree back to front (draw_tree, :2450-2517) with each leaf's bounds
// intersected with the part's viewport polygon, its polygons (draw_leaf, :2519-2792) and objects.
void emit_draw_list(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const Camera& cam, const ClippingPolygon& viewport,
	const int32_t clip_rect[4])
{
	draw_skybox(f, m, cam, viewport, cfg.textures_mip_bias, clip_rect);
	std::vector<uint32_t> stack;
	stack.push_back(uint32_t(m.nodes.size() - 1));
	const V4 w_row = cam.planes_matrix.row(3);
	while(!stack.empty())
	{
		const uint32_t node_index = stack.back();
		if(node_index >= FIRST_LEAF_INDEX)
		{
			stack.pop_back();
			const uint32_t leaf_index = node_index - FIRST_LEAF_INDEX;
			if(f.leaf_visible_frame[leaf_index] != f.frame_number)
				continue;
			ClippingPolygon lb = f.leaf_bounds[leaf_index];
			lb.intersect(viewport);
			if(!lb.valid())
				continue;
			V3 cp[8];
			lb.clip_planes(cp);
			const Leaf& leaf = m.leafs[leaf_index];
			for(uint32_t k = 0; k < leaf.num_polygons; ++k)
			{
				const uint32_t pi = leaf.first_polygon + k;
				const swh_frame::PolyData& pd = f.poly[pi];
				if(pd.visible_frame != f.frame_number)
					continue;
				const Polygon& p = m.polygons[pi];
				draw_polygon(f, cp, &f.vertices_transformed[p.first_vertex], p.num_vertices, pd.de, pd.te, pd.size, pd.surface_offset,
					SW_TEXSRC_SURFACE, m.materials[p.texture].blending_mode, clip_rect);
			}
			draw_leaf_objects(f, m, cfg, leaf_index, lb, clip_rect, cam.view_matrix, cam.planes_matrix);
		}
		else
		{
			const Node& node = m.nodes[node_index];
			const float w = dot(w_row, extend(node.plane.vec, -node.plane.dist));
			const uint32_t mask = w >= 0.0f ? 1u : 0u;
			stack.back() = node.children[1 ^ mask];
			stack.push_back(node.children[mask]);
		}
	}
	draw_view_models(f, m, viewport, clip_rect);
}

// ---------------------------------------------------------------------------------------------
// View portals: mirrors, cameras and parallax portals (partial_renderer.rs:1834-2282)
// ---------------------------------------------------------------------------------------------
void build_view(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const Camera& cam,
	const std::vector<uint32_t>* start_leafs, uint32_t portals_depth);

// prepare_view_portals_textures (partial_renderer.rs:1834-1996): each portal placed into the leaves its
// polygon reaches (position_portals, dynamic_objects_index.rs:144-157); a portal in a visible leaf that
// faces the camera and survives clipping by the union of its visible leaves' bounds gets a texture: its
// texture-space rect (from the world vertices, shrunk by 1/32) at the mip its screen projection asks
// for, and depth and texture-coordinate equations shifted to that rect. Textures are laid end to end
// in portal order.
void prepare_view_portals(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const Camera& cam)
{
	const uint32_t np = uint32_t(m.view_portals.size());
	f.portal_info.assign(np, swh_frame::PortalInfo{});
	f.portal_texels = 0;
	f.leaf_portals.clear();
	if(np == 0)
		return;
	f.leaf_portals.assign(m.leafs.size(), std::vector<uint32_t>());
	for(uint32_t pi = 0; pi < np; ++pi)
	{
		const ViewPortalDef& portal = m.view_portals[pi];
		swh_frame::PortalInfo& info = f.portal_info[pi];
		convex_hull_leafs_r(m, portal.vertices, 4, uint32_t(m.nodes.size() - 1), info.leafs);
		for(uint32_t l : info.leafs)
			f.leaf_portals[l].push_back(pi);
	}
	for(uint32_t pi = 0; pi < np; ++pi)
	{
		const ViewPortalDef& portal = m.view_portals[pi];
		swh_frame::PortalInfo& info = f.portal_info[pi];
		bool have_bounds = false;
		ClippingPolygon bounds;
		for(uint32_t l : info.leafs)
		{
			if(f.leaf_visible_frame[l] != f.frame_number)
				continue;
			if(have_bounds)
				bounds.extend(f.leaf_bounds[l]);
			else
			{
				bounds = f.leaf_bounds[l];
				have_bounds = true;
			}
		}
		if(!have_bounds)
			continue;
		V3 clip_planes[8];
		bounds.clip_planes(clip_planes);
		const V4 plane_transformed = mul(cam.planes_matrix, extend(portal.plane.vec, -portal.plane.dist));
		if(plane_transformed.w <= 0.0f)
			continue;
		V3 vt[4];
		for(int k = 0; k < 4; ++k)
			vt[k] = view_transform(cam.view_matrix, portal.vertices[k]);
		V2 v2d[MAX_VERTICES];
		const uint32_t n = project_and_clip_polygon(clip_planes, vt, 4, v2d);
		if(n < 3)
			continue;
		const float inf = float(1 << 29);
		float tc_min[2] = {inf, inf}, tc_max[2] = {-inf, -inf};
		for(int k = 0; k < 4; ++k)
			for(int i = 0; i < 2; ++i)
			{
				const float tc = dot(portal.tex_coord_equation[i].vec, portal.vertices[k]) + portal.tex_coord_equation[i].dist;
				if(tc < tc_min[i])
					tc_min[i] = tc;
				if(tc > tc_max[i])
					tc_max[i] = tc;
			}
		for(int i = 0; i < 2; ++i)
		{
			tc_min[i] += 1.0f / 32.0f;
			tc_max[i] -= 1.0f / 32.0f;
		}
		const DepthEq de = depth_from_plane(plane_transformed);
		const V4 tcb[2] = {mul(cam.planes_matrix, extend(portal.tex_coord_equation[0].vec, portal.tex_coord_equation[0].dist)),
			mul(cam.planes_matrix, extend(portal.tex_coord_equation[1].vec, portal.tex_coord_equation[1].dist))};
		const TcEq te = tc_from_depth_and_planes(de, tcb);
		const uint32_t mip = calculate_mip(v2d, n, de, te, cfg.textures_mip_bias);
		const float mip_scale = 1.0f / float(1 << mip);
		TcEq tes = scale(te, mip_scale);
		int32_t tmin[2], tmax[2];
		for(int i = 0; i < 2; ++i)
		{
			tmin[i] = saturating_f32_to_i32(std::floor(tc_min[i]) * mip_scale);
			tmax[i] = std::max(saturating_f32_to_i32(std::ceil(tc_max[i]) * mip_scale), tmin[i] + 1);
			const float t = float(tmin[i]);
			tes.dx[i] -= t * de.dx;
			tes.dy[i] -= t * de.dy;
			tes.k[i] -= t * de.k;
		}
		info.visible = true;
		info.resolution[0] = uint32_t(tmax[0] - tmin[0]);
		info.resolution[1] = uint32_t(tmax[1] - tmin[1]);
		info.texture_offset = f.portal_texels;
		info.de = de;
		info.te = tes;
		info.tc_min[0] = tmin[0];
		info.tc_min[1] = tmin[1];
		info.mip = mip;
		f.portal_texels += uint64_t(info.resolution[0]) * info.resolution[1];
		if(portal.texture_level != SW_NO_TEXTURE)
		{
			// partial_renderer.rs:2228-2258: the light at the centre of the portal's rect on its
			// plane (mip basis), scaled and offset; the texture's level at the portal's mip
			info.mix_level = portal.texture_level + mip;
			info.mix_blend = portal.texture_blending_mode;
			float light[3] = {0.0f, 0.0f, 0.0f};
			if(portal.texture_light_scale > 0.0f)
			{
				swh_frame::Basis b = form_basis(portal.plane, portal.tex_coord_equation);
				const float scale = float(1u << mip);
				b.u = b.u * scale;
				b.v = b.v * scale;
				const V3 centre = b.start + b.u * (float(tmin[0]) + float(info.resolution[0]) * 0.5f) +
					b.v * (float(tmin[1]) + float(info.resolution[1]) * 0.5f);
				grid_surface_light(m, centre, b.normal, light);
			}
			const float add[3] = {portal.texture_light_add.x, portal.texture_light_add.y, portal.texture_light_add.z};
			for(int i = 0; i < 3; ++i)
				info.mix_light[i] = light[i] * portal.texture_light_scale + add[i];
		}
	}
}

// The view matrix of a mirror or a parallax portal (partial_renderer.rs:2071-2176): world -> the
// portal's texture space at its mip, shifted to its rect, scaled so that the portal plane lies at
// Z_NEAR (all geometry on the near side is clipped).
M4 portal_view_matrix(const ViewPortalDef& portal, const swh_frame::PortalInfo& info, V3 eye, float dist_normalized)
{
	const float normal_len = length(portal.plane.vec);
	const float mip_scale = 1.0f / float(1 << info.mip);
	M4 tc_basis;
	const Plane* e = portal.tex_coord_equation;
	tc_basis.c[0] = v4(e[0].vec.x * mip_scale, e[0].vec.y * mip_scale, e[0].vec.z * mip_scale, e[0].dist * mip_scale);
	tc_basis.c[1] = v4(e[1].vec.x * mip_scale, e[1].vec.y * mip_scale, e[1].vec.z * mip_scale, e[1].dist * mip_scale);
	const float ns = mip_scale / normal_len;
	tc_basis.c[2] = v4(portal.plane.vec.x * ns, portal.plane.vec.y * ns, portal.plane.vec.z * ns, -portal.plane.dist * ns);
	tc_basis.c[3] = v4(0.0f, 0.0f, 0.0f, 1.0f);
	tc_basis = transpose(tc_basis);
	const V3 eye_tc = truncate(mul(tc_basis, extend(eye, 1.0f)));
	const M4 translate = from_translation(v3(-eye_tc.x, -eye_tc.y, -eye_tc.z));
	M4 shift_to_viewport = identity();
	shift_to_viewport.c[2].x = eye_tc.x - float(info.tc_min[0]);
	shift_to_viewport.c[2].y = eye_tc.y - float(info.tc_min[1]);
	const M4 z_scale = from_nonuniform_scale(Z_NEAR, Z_NEAR, Z_NEAR / (dist_normalized * mip_scale));
	return mul(mul(mul(shift_to_viewport, z_scale), translate), tc_basis);
}

Camera camera_from_view_matrix(V3 position, const M4& view)
{
	Camera c;
	c.position = position;
	c.view_matrix = view;
	if(!invert(transpose(view), c.planes_matrix))
		c.planes_matrix = identity();
	return c;
}

// build_view_portals_textures (partial_renderer.rs:2008-2200): every visible portal's view, rendered
// into its texture by the next level's renderer (a third-person view: no view models). Below the
// last level the textures stay black (zero).
void build_portal_views(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const Camera& cam, uint32_t depth)
{
	f.num_portal_views = 0;
	if(depth == 0)
		return;
	for(uint32_t pi = 0; pi < m.view_portals.size(); ++pi)
	{
		const swh_frame::PortalInfo& info = f.portal_info[pi];
		if(!info.visible)
			continue;
		const ViewPortalDef& portal = m.view_portals[pi];
		const float rw = float(info.resolution[0]), rh = float(info.resolution[1]);
		Camera pc;
		std::vector<uint32_t> transformed_leafs;
		const std::vector<uint32_t>* start = nullptr;
		if(portal.kind == PORTAL_CAMERA)
			pc = build_camera(portal.camera_position, portal.camera_yaw, portal.camera_pitch, portal.camera_fov, rw, rh);
		else if(portal.kind == PORTAL_MIRROR)
		{
			const float normal_square_len = dot(portal.plane.vec, portal.plane.vec);
			const float dist = dot(portal.plane.vec, cam.position) - portal.plane.dist;
			const V3 reflected = cam.position - portal.plane.vec * (2.0f * dist / normal_square_len);
			const float dist_normalized = dist / std::sqrt(normal_square_len);
			pc = camera_from_view_matrix(reflected, portal_view_matrix(portal, info, reflected, dist_normalized));
			start = &info.leafs; // the leaves the mirror lies in
		}
		else
		{
			V3 tv[4];
			for(int k = 0; k < 4; ++k)
				tv[k] = truncate(mul(portal.transform, extend(portal.vertices[k], 1.0f)));
			convex_hull_leafs_r(m, tv, 4, uint32_t(m.nodes.size() - 1), transformed_leafs);
			const float dist_normalized = (portal.plane.dist - dot(portal.plane.vec, cam.position)) / length(portal.plane.vec);
			M4 inv_transform;
			if(!invert(portal.transform, inv_transform))
				inv_transform = identity();
			const M4 view = mul(portal_view_matrix(portal, info, cam.position, dist_normalized), inv_transform);
			pc = camera_from_view_matrix(truncate(mul(portal.transform, extend(cam.position, 1.0f))), view);
			start = &transformed_leafs;
		}
		if(f.portal_views.size() <= f.num_portal_views)
			f.portal_views.emplace_back(new swh_frame());
		swh_frame& child = *f.portal_views[f.num_portal_views++];
		swh_frame_config cc = cfg;
		cc.width = info.resolution[0];
		cc.height = info.resolution[1];
		cc.view_model = 0;
		build_view(child, m, cc, pc, start, depth - 1);
	}
}

// prepare_frame + draw-list emission of one view (partial_renderer.rs:156-233, 1192-1283): visibility
// from the camera's leaf or from the given start leaves, lights, surfaces, inline models, models and
// decals, the portals and the views through them, then the draw list.
void build_view(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const Camera& cam,
	const std::vector<uint32_t>* start_leafs, uint32_t portals_depth)
{
	ensure_basis(f, m);
	f.frame_number++;
	f.visible_polygons.clear();
	f.surfaces.clear();
	f.lights.clear();
	f.light_indices.clear();
	f.draws.clear();
	f.vdata.clear();
	f.have_sky = false;
	f.num_triangles = 0;
	const float W = float(cfg.width), H = float(cfg.height);
	const ClippingPolygon frame_bounds = ClippingPolygon::from_box(0.0f, 0.0f, W, H);
	if(start_leafs)
		update_visibility_with_start_leafs(f, m, cam, frame_bounds, *start_leafs);
	else
		update_visibility(f, m, cam, frame_bounds);

	// dynamic lights and the leaves they touch
	build_dynamic_lights(f, m, cfg);

	// surfaces (partial_renderer.rs:1327-1452), leaf order
	uint64_t offset = 0;
	f.num_visible_leafs = 0;
	for(uint32_t li = 0; li < m.leafs.size(); ++li)
	{
		if(f.leaf_visible_frame[li] != f.frame_number)
			continue;
		f.num_visible_leafs++;
		V3 cp[8];
		f.leaf_bounds[li].clip_planes(cp);
		const Leaf& leaf = m.leafs[li];
		for(uint32_t k = 0; k < leaf.num_polygons; ++k)
			prepare_polygon_surface(f, m, cam, cp, offset, leaf.first_polygon + k, cfg.textures_mip_bias);
	}
	prepare_submodels(f, m, cfg, cam);
	f.first_submodel_surface = uint32_t(f.visible_polygons.size());
	prepare_submodel_surfaces(f, m, cfg, cam, offset);
	f.arena_texels = offset;
	for(uint32_t pi : f.visible_polygons)
		emit_surface(f, m, cfg, pi);

	// models and decals: animation, lighting, projection (per frame, before the tree walk)
	prepare_models(f, m, cfg, cam.view_matrix, cam.planes_matrix);
	prepare_decals(f, m, cfg, cam.planes_matrix);

	// portals and the views through them (after the surfaces, :229-233)
	prepare_view_portals(f, m, cfg, cam);
	build_portal_views(f, m, cfg, cam, portals_depth);

	// draw list: perform_rasterization (partial_renderer.rs:1192-1283). One thread: the whole screen
	// as clip rect and the screen box as viewport polygon. N threads: split_rect's N rects, each with
	// its integer clip rect and a viewport polygon widened by half a pixel; every rect gets the whole
	// traversal, so the draw list holds N runs of draws whose clip rects partition the screen.
	const uint32_t nthreads = cfg.virtual_threads <= 1 ? 1u : std::min<uint32_t>(cfg.virtual_threads, 64u);
	if(nthreads == 1)
	{
		const int32_t clip_rect[4] = {0, 0, int32_t(cfg.width), int32_t(cfg.height)};
		emit_draw_list(f, m, cfg, cam, frame_bounds, clip_rect);
	}
	else
	{
		std::vector<ScreenRect> rects;
		split_rect(ScreenRect{0.0f, 0.0f, W, H}, nthreads, rects);
		for(const ScreenRect& r : rects)
		{
			// `as i32` of the float bounds (:1250-1259), the polygon extended by 0.5 (:1261-1272)
			const int32_t clip_rect[4] = {int32_t(r.min_x), int32_t(r.min_y), int32_t(r.max_x), int32_t(r.max_y)};
			const ClippingPolygon viewport =
				ClippingPolygon::from_box(r.min_x - 0.5f, r.min_y - 0.5f, r.max_x + 0.5f, r.max_y + 0.5f);
			emit_draw_list(f, m, cfg, cam, viewport, clip_rect);
		}
	}

	sw_view& v = f.view;
	std::memset(&v, 0, sizeof(v));
	v.width = cfg.width;
	v.height = cfg.height;
	v.cam_pos[0] = cam.position.x;
	v.cam_pos[1] = cam.position.y;
	v.cam_pos[2] = cam.position.z;
	v.clear_background = (cfg.force_clear || (cfg.clear_background && !f.inside_leaf_volume)) ? 1u : 0u;
	// ColorVec::from_color_f32x3(&[32.0, 16.0, 8.0]).into() (partial_renderer.rs:255)
	v.background_color = cfg.color_bits == 64 ? 0x0000002000100008ull : 0x00201008ull;
	v.num_surfaces = uint32_t(f.surfaces.size());
	v.num_draws = uint32_t(f.draws.size());
}

} // namespace

extern "C" {

int swh_frame_create(swh_frame** out)
{
	if(!out)
		return -1;
	*out = new(std::nothrow) swh_frame();
	return *out ? 0 : -2;
}

void swh_frame_destroy(swh_frame* f) { delete f; }

int swh_frame_build(swh_frame* f, const swh_map* m, const swh_frame_config* cfg, const float cam_pos[3], float yaw, float pitch)
{
	if(!f || !m || !cfg || !cam_pos || cfg->width == 0 || cfg->height == 0 || (cfg->color_bits != 32 && cfg->color_bits != 64))
		return -1;
	const Camera cam = build_camera(v3(cam_pos[0], cam_pos[1], cam_pos[2]), yaw, pitch, cfg->fov, float(cfg->width), float(cfg->height));
	build_view(*f, *m, *cfg, cam, nullptr, std::min<uint32_t>(cfg->portals_depth, 4u));
	return 0;
}

int swh_frame_get(const swh_frame* f, swh_frame_data* out)
{
	if(!f || !out)
		return -1;
	out->view = f->view;
	out->surfaces = f->surfaces.data();
	out->num_surfaces = uint32_t(f->surfaces.size());
	out->lights = f->lights.data();
	out->num_lights = uint32_t(f->lights.size());
	out->light_indices = f->light_indices.data();
	out->num_light_indices = uint32_t(f->light_indices.size());
	out->draws = f->draws.data();
	out->num_draws = uint32_t(f->draws.size());
	out->vertex_data = f->vdata.data();
	out->num_vertex_ints = f->vdata.size();
	out->surface_arena_texels = f->arena_texels;
	out->num_visible_leafs = f->num_visible_leafs;
	out->num_triangles = f->num_triangles;
	out->first_submodel_surface = f->first_submodel_surface;
	out->num_portal_texels = f->portal_texels;
	out->num_portal_views = f->num_portal_views;
	out->num_visible_portals = 0;
	for(const swh_frame::PortalInfo& p : f->portal_info)
		out->num_visible_portals += p.visible ? 1u : 0u;
	return 0;
}

int swh_frame_portal_info(const swh_frame* f, uint32_t index, swh_portal_info* out)
{
	if(!f || !out || index >= f->portal_info.size())
		return -1;
	const swh_frame::PortalInfo& p = f->portal_info[index];
	std::memset(out, 0, sizeof(*out));
	out->visible = p.visible ? 1u : 0u;
	out->resolution[0] = p.resolution[0];
	out->resolution[1] = p.resolution[1];
	out->tc_min[0] = p.tc_min[0];
	out->tc_min[1] = p.tc_min[1];
	out->mip = p.mip;
	out->texture_offset = p.texture_offset;
	out->mix_level = p.visible ? p.mix_level : SW_NO_TEXTURE;
	out->mix_blend = p.mix_blend;
	for(int i = 0; i < 3; ++i)
		out->mix_light[i] = p.mix_light[i];
	return 0;
}

int swh_frame_model_data(const swh_frame* f, swh_model_data* out)
{
	if(!f || !out)
		return -1;
	const swh::ModelFrameData& md = f->models;
	std::memset(out, 0, sizeof(*out));
	out->instances = md.instances.data();
	out->num_instances = uint32_t(md.instances.size());
	out->runs = md.runs.data();
	out->num_runs = uint32_t(md.runs.size());
	out->run_positions = md.run_positions.data();
	out->run_num_draws = md.run_num_draws.data();
	out->prepared_vertices = md.prepared_vertices.data();
	out->num_prepared_vertices = md.prepared_vertices.size();
	out->prepared_triangles = md.prepared_triangles.data();
	out->num_prepared_triangles = md.prepared_triangles.size() / 3;
	out->prepared_counts = md.prepared_counts.data();
	out->num_prepared_counts = uint32_t(md.prepared_counts.size());
	out->vertex_capacity = md.vertex_capacity;
	out->triangle_capacity = md.triangle_capacity;
	return 0;
}

int64_t swh_assemble_model_draws(const sw_draw* draws, uint32_t num_draws, const sw_mesh_run* runs, const uint32_t* run_positions,
	uint32_t num_runs, const uint8_t* fans, uint64_t int_base, sw_draw* out, uint64_t capacity)
{
	if((num_draws && !draws) || (num_runs && (!runs || !run_positions || !fans)) || !out)
		return -1;
	uint64_t n = 0, ib = 0, fb = 0;
	uint32_t next = 0;
	auto put = [&](const sw_draw& d) {
		if(n >= capacity)
			return false;
		out[n++] = d;
		return true;
	};
	for(uint32_t r = 0; r < num_runs; ++r)
	{
		const sw_mesh_run& run = runs[r];
		if(run_positions[r] < next || run_positions[r] > num_draws)
			return -1;
		for(; next < run_positions[r]; ++next)
			if(!put(draws[next]))
				return -1;
		const uint64_t fan_max = 1u + run.num_planes3d + run.num_planes2d;
		sw_draw d;
		std::memset(&d, 0, sizeof(d));
		d.kind = SW_DRAW_TRIANGLE;
		d.blend = run.blend;
		d.num_vertices = 3;
		d.texture_offset = run.texture_offset;
		d.texture_size[0] = run.texture_size[0];
		d.texture_size[1] = run.texture_size[1];
		for(int i = 0; i < 4; ++i)
			d.clip_rect[i] = run.clip_rect[i];
		for(uint32_t t = 0; t < run.max_triangles; ++t)
			for(uint32_t j = 0; j < fans[fb + t]; ++j)
			{
				const uint64_t first = int_base + ib + (uint64_t(t) * fan_max + j) * 21u;
				if(first > 0xFFFFFFFFull)
					return -1;
				d.first_vertex_int = uint32_t(first);
				if(!put(d))
					return -1;
			}
		ib += uint64_t(run.max_triangles) * fan_max * 21u;
		fb += run.max_triangles;
	}
	for(; next < num_draws; ++next)
		if(!put(draws[next]))
			return -1;
	return int64_t(n);
}

int swh_frame_portal_view(const swh_frame* f, uint32_t index, const swh_frame** out)
{
	if(!f || !out || index >= f->num_portal_views)
		return -1;
	*out = f->portal_views[index].get();
	return 0;
}

// ---------------------------------------------------------------------------------------------
// Shadow-map draw lists: DepthRenderer::draw_map (depth_renderer.rs:47-287) for the six sides of a
// cube light (light.rs:73-86, partial_renderer.rs:541-566). The map's static geometry, back to front,
// visible leaves only; the pixel work is sw_render_depth's.
// ---------------------------------------------------------------------------------------------
int swh_shadow_create(swh_shadow** out)
{
	if(!out)
		return -1;
	*out = new(std::nothrow) swh_shadow();
	return *out ? 0 : -1;
}

void swh_shadow_destroy(swh_shadow* s) { delete s; }

void swh_shadow_clear(swh_shadow* s)
{
	if(!s)
		return;
	s->targets.clear();
	s->draws.clear();
	s->vdata.clear();
	s->texels = 0;
}

// One shadow map: DepthRenderer::draw_map (depth_renderer.rs:47-109, 202-287) from the camera `cam`,
// appended as a size x size target at the running texel offset.
static void add_shadow_target(swh_shadow* s, const swh_map* m, const Camera& cam, uint32_t size, uint32_t flags)
{
	const float fs = float(size);
	const ClippingPolygon frame_bounds = ClippingPolygon::from_box(0.0f, 0.0f, fs, fs);
	swh_frame& vis = s->vis;
	vis.frame_number++;
	update_visibility(vis, *m, cam, frame_bounds);
	sw_depth_target t;
	t.offset = s->texels;
	t.width = size;
	t.height = size;
	t.first_draw = uint32_t(s->draws.size());
	t.flags = flags;
	t.reserved = 0;
	// draw_tree (:69-109): back to front, leaves with visibility bounds
	std::vector<uint32_t> stack;
	stack.push_back(uint32_t(m->nodes.size() - 1));
	const V4 w_row = cam.planes_matrix.row(3);
	while(!stack.empty())
	{
		const uint32_t node_index = stack.back();
		if(node_index >= FIRST_LEAF_INDEX)
		{
			stack.pop_back();
			const uint32_t leaf_index = node_index - FIRST_LEAF_INDEX;
			if(vis.leaf_visible_frame[leaf_index] != vis.frame_number)
				continue;
			V3 cp[8];
			vis.leaf_bounds[leaf_index].clip_planes(cp);
			const Leaf& leaf = m->leafs[leaf_index];
			for(uint32_t k = 0; k < leaf.num_polygons; ++k)
			{
				// draw_polygon (:202-254)
				const Polygon& p = m->polygons[leaf.first_polygon + k];
				// materials_shadow_table (:23-37): blended materials cast no shadows
				if(m->materials[p.texture].blending_mode != SW_BLEND_NONE)
					continue;
				const V4 pt = mul(cam.planes_matrix, extend(p.plane.vec, -p.plane.dist));
				if(pt.w <= 0.0f)
					continue; // back face
				const float w = -pt.w;
				const float dx = pt.x / w, dy = pt.y / w;
				const float bias_const = -1.0f / float(1 << 20), bias_slope = -1.0f, eq_scale = 1.0f - 1.0f / 1024.0f;
				sw_depth_draw d;
				d.depth_equation[0] = eq_scale * dx;
				d.depth_equation[1] = eq_scale * dy;
				d.depth_equation[2] = eq_scale * (pt.z / w + bias_const + bias_slope * (std::fabs(dx) + std::fabs(dy)));
				d.depth_test = 0;
				const uint32_t n = std::min(p.num_vertices, MAX_VERTICES);
				if(n < 3)
					continue;
				V3 vt[MAX_VERTICES];
				for(uint32_t i = 0; i < n; ++i)
					vt[i] = view_transform(cam.view_matrix, m->vertices[p.first_vertex + i]);
				// draw_depth_polygon (:257-287)
				V2 v2d[MAX_VERTICES];
				const uint32_t nc = project_and_clip_polygon(cp, vt, n, v2d);
				if(nc < 3)
					continue;
				d.first_vertex_int = uint32_t(s->vdata.size());
				d.num_vertices = nc;
				for(uint32_t i = 0; i < nc; ++i)
				{
					s->vdata.push_back(saturating_f32_to_i32(v2d[i].x * 65536.0f));
					s->vdata.push_back(saturating_f32_to_i32(v2d[i].y * 65536.0f));
				}
				s->draws.push_back(d);
			}
		}
		else
		{
			const Node& node = m->nodes[node_index];
			const float wn = dot(w_row, extend(node.plane.vec, -node.plane.dist));
			const uint32_t mask = wn >= 0.0f ? 1u : 0u;
			stack.back() = node.children[1 ^ mask];
			stack.push_back(node.children[mask]);
		}
	}
	t.num_draws = uint32_t(s->draws.size()) - t.first_draw;
	s->targets.push_back(t);
	s->texels += uint64_t(size) * size;
}

int swh_shadow_add_cube(swh_shadow* s, const swh_map* m, const float position[3], uint32_t size)
{
	if(!s || !m || !position || size == 0 || size > 65535)
		return -1;
	ensure_basis(s->vis, *m);
	const V3 pos = v3(position[0], position[1], position[2]);
	const float fs = float(size);
	for(uint32_t side = 0; side < 6; ++side)
	{
		add_shadow_target(s, m, complete_view(pos, cube_side_rotation(side), v2(1.0f, 1.0f), fs, fs), size, 0);
	}
	return 0;
}

// PartialRenderer::build_shadow_maps' projector branch (partial_renderer.rs:567-584): the camera of
// calculate_projector_shadow_map_matrices (light.rs:63-71, matrix.rs:31-80), the map drawn, then
// make_shadow_map_circle (the target's SW_DEPTH_CIRCLE). basis_out: the sw_light_shadow basis of
// create_dynamic_light_projector_shadow_map (renderer_structs.rs:288-307), x, y, z.
int swh_shadow_add_projector(swh_shadow* s, const swh_map* m, const float position[3], const float rotation[4], float fov,
	uint32_t size, float basis_out[9])
{
	if(!s || !m || !position || !rotation || size == 0 || size > 65535 || !(fov > 0.0f && fov < 3.14159265f))
		return -1;
	ensure_basis(s->vis, *m);
	const V3 pos = v3(position[0], position[1], position[2]);
	const Quat q = Quat{rotation[0], v3(rotation[1], rotation[2], rotation[3])};
	const float fs = float(size);
	const float inv_half_fov_tan = 1.0f / std::tan(fov * 0.5f);
	const V2 fov_scale = v2(inv_half_fov_tan / (fs / fs), inv_half_fov_tan);
	M4 basis_change = identity();
	basis_change.c[0].x = 0.0f;
	basis_change.c[1].y = 0.0f;
	basis_change.c[2].z = 0.0f;
	basis_change.c[0].z = 1.0f;
	basis_change.c[1].x = -1.0f;
	basis_change.c[2].y = -1.0f;
	const Quat conj = Quat{q.s, v3(-q.v.x, -q.v.y, -q.v.z)};
	add_shadow_target(s, m, complete_view(pos, mul(basis_change, quat_to_m4(conj)), fov_scale, fs, fs), size, SW_DEPTH_CIRCLE);
	if(basis_out)
	{
		// cgmath's Quaternion::rotate_vector: v + 2 q.v x (q.v x v + s v)
		auto rotate = [&](V3 v) {
			const V3 tmp = cross(q.v, v) + v * q.s;
			return cross(q.v, tmp) * 2.0f + v;
		};
		const V3 bx = rotate(v3(0.0f, 1.0f, 0.0f)) * inv_half_fov_tan;
		const V3 by = rotate(v3(0.0f, 0.0f, 1.0f)) * inv_half_fov_tan;
		const V3 bz = rotate(v3(-1.0f, -0.0f, -0.0f));
		const float b[9] = {bx.x, bx.y, bx.z, by.x, by.y, by.z, bz.x, bz.y, bz.z};
		std::memcpy(basis_out, b, sizeof(b));
	}
	return 0;
}

int swh_shadow_get(const swh_shadow* s, swh_shadow_data* out)
{
	if(!s || !out)
		return -1;
	out->targets = s->targets.data();
	out->num_targets = uint32_t(s->targets.size());
	out->draws = s->draws.data();
	out->num_draws = uint32_t(s->draws.size());
	out->vertex_data = s->vdata.data();
	out->num_vertex_ints = s->vdata.size();
	out->num_depth_texels = s->texels;
	return 0;
}

} // extern "C"
