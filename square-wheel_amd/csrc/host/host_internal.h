// Internal data model of the CPU front-end (libsw_host.so).
//
// Map layout follows common/bsp_map_compact.rs:7-199 (BSPNode / BSPLeaf / Polygon / Portal,
// FIRST_LEAF_INDEX = 1 << 31, root node last, lightmap offset 0 = "no lightmap"); math follows
// cgmath 0.18's conventions as used by the reference (column-major Matrix4, Vector*Matrix via row
// dot products).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../../include/sw_host.h"

namespace swh
{

struct V2
{
	float x, y;
};
struct V3
{
	float x, y, z;
};
struct V4
{
	float x, y, z, w;
};

inline V2 v2(float x, float y) { return V2{x, y}; }
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V4 v4(float x, float y, float z, float w) { return V4{x, y, z, w}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
inline V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
inline V2 operator*(V2 a, float s) { return v2(a.x * s, a.y * s); }
inline V2 operator/(V2 a, float s) { return v2(a.x / s, a.y / s); }
// cgmath dot: element-wise product, then sum left to right
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline float dot(V4 a, V4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
inline V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V4 extend(V3 a, float w) { return v4(a.x, a.y, a.z, w); }
inline V3 truncate(V4 a) { return v3(a.x, a.y, a.z); }

// Column-major 4x4 (cgmath Matrix4): c[i] is column i.
struct M4
{
	V4 c[4];
	float& at(int row, int col) { return (&c[col].x)[row]; }
	float at(int row, int col) const { return (&c[col].x)[row]; }
	V4 row(int r) const { return v4(at(r, 0), at(r, 1), at(r, 2), at(r, 3)); }
};

M4 identity();
M4 mul(const M4& a, const M4& b);
V4 mul(const M4& m, V4 v);
M4 transpose(const M4& m);
bool invert(const M4& m, M4& out);
M4 from_nonuniform_scale(float x, float y, float z);
M4 from_translation(V3 t);

struct Plane
{
	V3 vec;
	float dist;
};

struct Polygon
{
	uint32_t first_vertex, num_vertices;
	Plane plane;
	Plane tex_coord_equation[2];
	int32_t tex_coord_min[2], tex_coord_max[2];
	uint32_t lightmap_data_offset;
	uint32_t texture; // material index
};

struct Node
{
	uint32_t children[2]; // >= FIRST_LEAF_INDEX: leaf
	Plane plane;
};

struct Leaf
{
	uint32_t first_polygon, num_polygons, first_leaf_portal, num_leaf_portals;
};

// Inline (brush) models (bsp_map_compact.rs Submodel / SubmodelBSPNode): a submodel's polygons are
// a range of the polygon array after the leafs' polygons; its own BSP orders them. A child index at
// or past the node count means "no child".
struct Submodel
{
	uint32_t root_node, first_polygon, num_polygons;
};

struct SubmodelNode
{
	Plane plane;
	uint32_t first_polygon, num_polygons;
	uint32_t children[2]; // [0]: front of the plane, [1]: back
};

const uint32_t NO_POLYGON_LEAF = 0xFFFFFFFFu;

// A view portal (frame_info.rs:171-196 ViewPortal): a polygon showing another view, placed by the
// game in the reference and seeded with the map here. Mirrors reflect the camera in the portal's
// plane; a camera portal shows a camera at a fixed position; a parallax portal moves the camera by
// a fixed transform.
enum : uint32_t
{
	PORTAL_MIRROR = 0,
	PORTAL_CAMERA = 1,
	PORTAL_PARALLAX = 2,
};

struct ViewPortalDef
{
	uint32_t kind;
	Plane plane; // facing the side the portal is seen from
	Plane tex_coord_equation[2];
	V3 vertices[4];
	uint32_t blending_mode;
	V3 camera_position; // PORTAL_CAMERA
	float camera_yaw, camera_pitch, camera_fov;
	M4 transform; // PORTAL_PARALLAX: camera position -> the portal's view point
	// ViewPortalTexture (frame_info.rs:198-205): an overlay mixed over the portal texture, or none
	uint32_t texture_level = SW_NO_TEXTURE; // lite level of the texture's mip 0 (its mips follow)
	uint32_t texture_blending_mode = SW_BLEND_AVERAGE;
	float texture_light_scale = 1.0f;
	V3 texture_light_add{0.0f, 0.0f, 0.0f};
};

struct Portal
{
	uint32_t leafs[2];
	Plane plane;
	uint32_t first_vertex, num_vertices;
};

const uint32_t FIRST_LEAF_INDEX = 1u << 31;
const uint32_t LIGHTMAP_SCALE_LOG2 = 4;
const uint32_t LIGHTMAP_SCALE = 1u << LIGHTMAP_SCALE_LOG2;
const uint32_t MAX_MIP = 4;

struct Material
{
	uint32_t first_level; // index of mip 0 in the level table; mips follow
	float roughness;
	bool is_metal;
	uint32_t blending_mode;
	bool sky;
};

struct BBox
{
	V3 min, max;
};

struct PointLight
{
	V3 pos;
	float intensity;
	V3 color;
};

// LightGridElement (bsp_map_compact.rs): a light cube (+x, -x... order of light_cube.rs:24-64) plus
// one dominant directional light.
struct LightGridElement
{
	float light_cube[6][3];
	V3 light_direction_vector_scaled;
	float directional_light_color[3];
};

// LightCube (common/light_cube.rs:3-160): accumulates light samples by direction into six
// signed axis bins plus a per-channel sum of scaled directions, then splits off one dominant
// directional light.
struct LightCubeAccum
{
	float cube[6][3] = {};
	V3 sum_dirs[3] = {v3(0, 0, 0), v3(0, 0, 0), v3(0, 0, 0)};

	void add_sample(V3 direction, const float light[3])
	{
		const float len = std::max(length(direction), 0.00000001f);
		const V3 d = direction / len;
		const float comps[3] = {d.x, d.y, d.z};
		for(int axis = 0; axis < 3; ++axis)
		{
			const int side = comps[axis] <= 0.0f ? axis * 2 : axis * 2 + 1;
			const float k = comps[axis] <= 0.0f ? -comps[axis] : comps[axis];
			for(int i = 0; i < 3; ++i)
				cube[side][i] += light[i] * k;
		}
		for(int i = 0; i < 3; ++i)
			sum_dirs[i] = sum_dirs[i] + d * light[i];
	}
	void add_constant(const float light[3])
	{
		for(int side = 0; side < 6; ++side)
			for(int i = 0; i < 3; ++i)
				cube[side][i] += light[i];
	}
	void scale(float k)
	{
		for(int side = 0; side < 6; ++side)
			for(int i = 0; i < 3; ++i)
				cube[side][i] *= k;
		for(int i = 0; i < 3; ++i)
			sum_dirs[i] = sum_dirs[i] * k;
	}
	LightGridElement convert() const
	{
		LightGridElement e{};
		const V3 dir = (sum_dirs[0] + sum_dirs[1] + sum_dirs[2]) * 0.33f;
		const float len2 = std::max(dot(dir, dir), 0.00000001f);
		float color[3];
		for(int i = 0; i < 3; ++i)
			color[i] = std::max(dot(sum_dirs[i], dir), 0.0f) / len2;
		for(int side = 0; side < 6; ++side)
			for(int i = 0; i < 3; ++i)
				e.light_cube[side][i] = cube[side][i];
		const float comps[3] = {dir.x, dir.y, dir.z};
		for(int i = 0; i < 3; ++i)
			for(int axis = 0; axis < 3; ++axis)
			{
				const int side = comps[axis] <= 0.0f ? axis * 2 : axis * 2 + 1;
				const float k = comps[axis] <= 0.0f ? -comps[axis] : comps[axis];
				e.light_cube[side][i] = std::max(e.light_cube[side][i] - color[i] * k, 0.0f);
			}
		e.light_direction_vector_scaled = dir;
		for(int i = 0; i < 3; ++i)
			e.directional_light_color[i] = color[i];
		return e;
	}
};

// A vertex-animated triangle model (triangle_model.rs VertexData::VertexAnimated): positions and
// normals per frame, texture coordinates shared, in model space.
struct ModelMesh
{
	uint32_t num_vertices = 0, num_frames = 0;
	std::vector<V3> pos; // [frame][vertex]
	std::vector<V3> normal; // [frame][vertex]
	std::vector<float> tc; // 2 per vertex, in [0, 1]
	std::vector<uint32_t> tris; // 3 per triangle
	std::vector<BBox> frame_bbox;
	// sw_model_mesh's VERTEX_ANIMATED layout: {position, normal} per vertex per frame, and the
	// triangles in the winding the camera sees front faces clockwise
	std::vector<float> abi_frame_vertices;
	std::vector<uint16_t> abi_triangles;
};

} // namespace swh

struct swh_map
{
	swh_map_config config;
	uint32_t gx, gy, gz;
	std::vector<uint8_t> cell; // 0 empty, 1 solid, 2 sky
	std::vector<uint32_t> cell_leaf;
	std::vector<swh::BBox> leaf_bbox;
	std::vector<uint32_t> empty_cells;

	std::vector<swh::Node> nodes;
	std::vector<swh::Leaf> leafs;
	std::vector<swh::Polygon> polygons;
	std::vector<swh::Portal> portals;
	std::vector<uint32_t> leafs_portals;
	std::vector<swh::V3> vertices;
	std::vector<swh::Material> materials;
	std::vector<uint32_t> polygon_leaf; // NO_POLYGON_LEAF for submodel polygons
	std::vector<swh::Submodel> submodels;
	std::vector<swh::SubmodelNode> submodel_nodes;
	std::vector<uint32_t> submodel_cell; // the empty cell each generated submodel stands in
	std::vector<swh::ViewPortalDef> view_portals;

	std::vector<float> lm_simple; // 3 floats per element
	std::vector<sw_directional_lightmap_element> lm_dir;
	std::vector<sw_texel> texels;
	std::vector<uint32_t> lite;
	std::vector<sw_texture_level> levels;
	std::vector<uint32_t> decal_levels; // mip-0 level of each decal texture
	std::vector<uint32_t> model_levels; // mip-0 level of each model skin
	std::vector<uint32_t> sky32;
	std::vector<uint64_t> sky64;
	uint32_t sky_side_size = 0;
	std::vector<swh::PointLight> lights;
	swh::V3 sun_dir{0, 0, -1};
	float sun_intensity = 0.0f;
	// light grid: one sample at the centre of every cell (zero for solid cells)
	std::vector<swh::LightGridElement> light_grid;
	swh::ModelMesh model_mesh;
};

namespace swh
{
inline uint32_t cell_index(const swh_map& m, uint32_t x, uint32_t y, uint32_t z) { return (z * m.gy + y) * m.gx + x; }
uint32_t find_leaf(const swh_map& m, V3 p);
} // namespace swh
