// Frame state shared by the front-end's translation units.
#pragma once

#include <memory>

#include "host_internal.h"
#include "host_models.h"

namespace swh
{

const uint32_t MAX_VERTICES = 24; // renderer_utils.rs:4
const float Z_NEAR = 1.0f; // renderer_utils.rs:7

// ---------------------------------------------------------------------------------------------
// ClippingPolygon (clipping_polygon.rs)
// ---------------------------------------------------------------------------------------------
struct ClipAxis
{
	float min, max;
	bool valid() const { return min < max; }
	bool contains(const ClipAxis& o) const { return o.min >= min && o.max <= max; }
	void extend(const ClipAxis& o)
	{
		if(o.min < min)
			min = o.min;
		if(o.max > max)
			max = o.max;
	}
	void extend_point(float p)
	{
		if(p < min)
			min = p;
		if(p > max)
			max = p;
	}
	void intersect(const ClipAxis& o)
	{
		if(o.min > min)
			min = o.min;
		if(o.max < max)
			max = o.max;
	}
};

struct ClippingPolygon
{
	ClipAxis x, y, xpy, xmy;
	static ClippingPolygon from_box(float min_x, float min_y, float max_x, float max_y)
	{
		return ClippingPolygon{{min_x, max_x}, {min_y, max_y}, {min_x + min_y, max_x + max_y}, {min_x - max_y, max_x - min_y}};
	}
	static ClippingPolygon from_point(V2 p)
	{
		return ClippingPolygon{{p.x, p.x}, {p.y, p.y}, {p.x + p.y, p.x + p.y}, {p.x - p.y, p.x - p.y}};
	}
	bool valid() const { return x.valid() && y.valid() && xpy.valid() && xmy.valid(); }
	bool contains(const ClippingPolygon& o) const
	{
		return x.contains(o.x) && y.contains(o.y) && xpy.contains(o.xpy) && xmy.contains(o.xmy);
	}
	void extend(const ClippingPolygon& o)
	{
		x.extend(o.x);
		y.extend(o.y);
		xpy.extend(o.xpy);
		xmy.extend(o.xmy);
	}
	void extend_point(V2 p)
	{
		x.extend_point(p.x);
		y.extend_point(p.y);
		xpy.extend_point(p.x + p.y);
		xmy.extend_point(p.x - p.y);
	}
	void intersect(const ClippingPolygon& o)
	{
		x.intersect(o.x);
		y.intersect(o.y);
		xpy.intersect(o.xpy);
		xmy.intersect(o.xmy);
	}
	void clip_planes(V3 out[8]) const
	{
		out[0] = v3(-1.0f, 0.0f, -x.max);
		out[1] = v3(1.0f, 0.0f, x.min);
		out[2] = v3(0.0f, -1.0f, -y.max);
		out[3] = v3(0.0f, 1.0f, y.min);
		out[4] = v3(-1.0f, -1.0f, -xpy.max);
		out[5] = v3(1.0f, 1.0f, xpy.min);
		out[6] = v3(-1.0f, 1.0f, -xmy.max);
		out[7] = v3(1.0f, -1.0f, xmy.min);
	}
};

struct DepthEq
{
	float dx, dy, k;
	float sample(V2 p) const { return dx * p.x + dy * p.y + k; }
};
struct TcEq
{
	float dx[2], dy[2], k[2];
};


struct Camera
{
	V3 position;
	M4 view_matrix;
	M4 planes_matrix;
};

uint32_t clip_3d_polygon_by_z_plane(const V3* poly, uint32_t n, float z_dist, V3* out, uint32_t cap);
V3 line_z_intersection(V3 v0, V3 v1, float z);
DepthEq depth_from_plane(V4 p);
TcEq tc_from_depth_and_planes(const DepthEq& d, const V4 t[2]);
TcEq scale(const TcEq& e, float s);
bool affine_may_be_used(const DepthEq& de, const TcEq& te, V2 min_p, V2 max_p);
uint32_t calculate_mip(const V2* points, uint32_t n, const DepthEq& de, const TcEq& te, float mip_bias);
int32_t saturating_f32_to_i32(float x);
uint32_t clip_2d_polygon(const V2* poly, uint32_t n, V3 line, V2* out, uint32_t cap);
uint32_t project_and_clip_polygon(const V3 clip_planes[8], const V3* verts, uint32_t n, V2* out);
V3 view_transform(const M4& m, V3 v);
int32_t f32_to_fixed16_unchecked(float x);
void draw_polygon(swh_frame& f, const V3 clip_planes[8], const V3* verts, uint32_t nverts, const DepthEq& de, const TcEq& te,
	const uint32_t tsize[2], uint64_t texture_offset, uint32_t texture_source, uint32_t blend, const int32_t clip_rect[4]);
// Decals of the leaf's polygons, then the leaf's models, after the leaf's polygons (draw_leaf,
// partial_renderer.rs:2519-2792). `bounds` is the leaf's screen ClippingPolygon.
void draw_leaf_objects(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, uint32_t leaf, const ClippingPolygon& bounds,
	const int32_t clip_rect[4], const M4& view, const M4& planes);
void draw_view_models(swh_frame& f, const swh_map& m, const ClippingPolygon& viewport, const int32_t clip_rect[4]);

} // namespace swh

using swh::ClippingPolygon;
using swh::DepthEq;
using swh::ModelFrameData;
using swh::TcEq;
using swh::V3;

// =============================================================================================
// Frame state
// =============================================================================================
struct swh_frame
{
	const swh_map* basis_map = nullptr;
	struct Basis
	{
		V3 u, v, start, normal;
	};
	std::vector<Basis> basis;

	struct PolyData
	{
		uint32_t visible_frame = 0;
		DepthEq de;
		TcEq te;
		uint64_t surface_offset;
		uint32_t size[2];
		uint32_t mip;
		int32_t tc_min[2];
	};
	std::vector<PolyData> poly;
	uint32_t frame_number = 0;
	std::vector<V3> vertices_transformed;

	// visibility
	std::vector<uint32_t> leaf_visible_frame;
	std::vector<ClippingPolygon> leaf_bounds;
	std::vector<uint32_t> portal_visible_frame;
	std::vector<uint8_t> portal_has_projection;
	std::vector<ClippingPolygon> portal_projection;
	bool inside_leaf_volume = true;

	bool have_sky = false;
	uint32_t sky_material = 0;
	ClippingPolygon sky_bounds;

	// outputs
	std::vector<uint32_t> visible_polygons;
	std::vector<sw_surface_desc> surfaces;
	std::vector<sw_light> lights;
	std::vector<uint32_t> light_indices;
	std::vector<sw_draw> draws;
	std::vector<int32_t> vdata;
	uint64_t arena_texels = 0;
	sw_view view{};
	uint32_t num_visible_leafs = 0;
	uint32_t num_triangles = 0;
	uint32_t first_submodel_surface = 0; // surfaces from here on belong to inline models

	// dynamic objects
	std::vector<std::vector<uint32_t>> leaf_lights;
	ModelFrameData models;

	// inline models of the frame (InlineModelsIndex + SubmodelInfo, partial_renderer.rs:593-664)
	struct SubmodelFrame
	{
		bool active = false; // placed by the frame's entity
		swh::BBox bbox; // initial bounding box of the polygons, extended by 1/4 (model space)
		swh::M4 model; // model matrix: translation(position) * rotation * translation(-centre)
		swh::M4 world_planes; // inverse of the model matrix's transpose: transforms planes
		swh::M4 view, planes; // the camera's matrices combined with the model's
		std::vector<uint32_t> leafs; // leaves the transformed bbox touches
		std::vector<uint32_t> lights; // visible dynamic lights whose sphere's box touches the bbox
	};
	std::vector<SubmodelFrame> submodels;
	std::vector<std::vector<uint32_t>> leaf_submodels; // per leaf: indices into `submodels`

	// view portals of the frame (PortalInfo, partial_renderer.rs:1834-1997)
	struct PortalInfo
	{
		bool visible = false; // resolution > 0
		uint32_t resolution[2] = {0, 0};
		uint64_t texture_offset = 0; // into the frame's portal textures (ColorT texels)
		DepthEq de;
		TcEq te;
		int32_t tc_min[2] = {0, 0};
		uint32_t mip = 0;
		std::vector<uint32_t> leafs; // the leaves the portal polygon's convex hull reaches
		// the overlay of a portal with a texture (partial_renderer.rs:2205-2276): the texture's
		// level at the portal's mip, and the light at the portal's centre
		uint32_t mix_level = SW_NO_TEXTURE;
		uint32_t mix_blend = SW_BLEND_NONE;
		float mix_light[3] = {0.0f, 0.0f, 0.0f};
	};
	std::vector<PortalInfo> portal_info;
	std::vector<std::vector<uint32_t>> leaf_portals; // per leaf: portal indices
	uint64_t portal_texels = 0; // sum of the portal textures' areas
	// the views seen through the visible portals, one frame each, in portal order (when the portals
	// depth allows another level): their framebuffers laid end to end are the portal textures
	std::vector<std::unique_ptr<swh_frame>> portal_views;
	uint32_t num_portal_views = 0;
};

// Shadow-map draw lists under construction: the visibility state of the map (leaf bounds) reused
// side after side, and the depth targets, draws and Fixed16 vertex data of sw_render_depth.
struct swh_shadow
{
	swh_frame vis;
	std::vector<sw_depth_target> targets;
	std::vector<sw_depth_draw> draws;
	std::vector<int32_t> vdata;
	uint64_t texels = 0;
};


