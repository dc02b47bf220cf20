// Dynamic objects of a frame: dynamic point lights, animated triangle models and decals.
#pragma once

#include "host_internal.h"

struct swh_frame;

namespace swh
{

// ModelVertex3d (triangle_models_rendering.rs): camera-space position, texture coordinates in
// texels, light.
struct ModelVertex3d
{
	V3 pos;
	float tc[2];
	float light[3];
};

// One model entity of the frame (frame_info.rs ModelEntity, with VertexAnimated data and
// ModelLighting::Default).
struct ModelEntity
{
	V3 pos;
	float yaw;
	uint32_t frames[2];
	float lerp; // weight of frames[0]
	uint32_t skin; // mip-0 lite level of its texture
	uint32_t blend;
	bool view_model = false; // ModelEntityDrawFlags::VIEW_MODEL: drawn after the tree, over the viewport
};

// One visible model with its transformed vertices and its back-to-front sorted front-facing
// triangles (partial_renderer.rs VisibleDynamicMeshInfo).
struct VisibleModel
{
	uint32_t entity;
	uint32_t first_vertex, num_vertices;
	uint32_t first_triangle, num_triangles;
	uint32_t level; // lite texture level of the selected mip
	V3 bbox_vertices[8]; // camera space
	float clip[4][2]; // screen ClippingPolygon axes (x, y, x + y, x - y) as min/max pairs
	float depth; // camera-space depth of the bbox centre
	BBox bbox; // model space: the union of the two animation frames' boxes
	M4 view, planes; // the camera's matrices combined with the model matrix (ordering within a leaf)
};

// Decal (frame_info.rs:113-123): a unit cube (half-size 1) scaled and rotated, projected onto
// the polygons it touches.
struct DecalEntity
{
	V3 pos;
	V3 axis[3]; // rotation columns (unit), decal-space x, y, z in world space
	V3 scale;
	uint32_t level; // mip-0 lite level
	uint32_t blend;
	float lightmap_light_scale;
	float light_add[3];
};

struct ModelFrameData
{
	std::vector<ModelEntity> entities;
	std::vector<VisibleModel> visible;
	std::vector<ModelVertex3d> vertices;
	std::vector<uint32_t> triangles; // 3 per triangle
	std::vector<std::vector<uint32_t>> leaf_models; // per leaf: indices into `visible`
	std::vector<uint32_t> view_models; // indices into `visible`
	std::vector<DecalEntity> decals;
	struct DecalInfo
	{
		M4 camera_planes_matrix;
		LightGridElement dynamic_light;
	};
	std::vector<DecalInfo> decal_info;
	std::vector<std::vector<uint32_t>> leaf_decals; // per leaf: decal indices
	// the models as instances of sw_prepare_models and the meshes drawn in leaves as runs
	bool device = false; // swh_frame_config::device_models: no host animation, runs instead of triangles
	std::vector<sw_model_instance> instances; // per visible model
	std::vector<sw_mesh_run> runs;
	std::vector<uint32_t> run_positions, run_num_draws;
	std::vector<sw_model_vertex> prepared_vertices; // host preparation: `vertices` in the ABI layout
	std::vector<uint16_t> prepared_triangles;
	std::vector<uint32_t> prepared_counts;
	uint64_t vertex_capacity = 0, triangle_capacity = 0;
};

void build_dynamic_lights(swh_frame& f, const swh_map& m, const swh_frame_config& cfg);
void prepare_models(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const M4& view, const M4& planes);
void prepare_decals(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const M4& planes);
// The light-grid light of a surface point with normal n (partial_renderer.rs:2232-2247):
// get_light_cube_light of the grid's cube plus the directional light by max(dot(n, dir), 0).
void grid_surface_light(const swh_map& m, V3 pos, V3 n, float out[3]);

} // namespace swh
