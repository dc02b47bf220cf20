// Dynamic objects of a frame: dynamic point lights, animated triangle models and decals.
#pragma once

#include "host_internal.h"

struct swh_frame;

namespace swh
{

struct ModelTriangle
{
	uint32_t v[3];
};

struct ModelVertex3d
{
	V3 pos; // camera space (view_matrix applied, not projected)
	float tc[2];
	float light[3];
};

// One visible model of the frame, with its transformed vertices and its back-to-front sorted
// front-facing triangles (triangle_models_rendering.rs:324-391).
struct VisibleModel
{
	uint32_t model;
	uint32_t first_vertex, num_vertices;
	uint32_t first_triangle, num_triangles;
	uint32_t skin_level;
	BBox bbox;
	std::vector<uint32_t> leafs;
};

struct ModelFrameData
{
	std::vector<VisibleModel> visible;
	std::vector<ModelVertex3d> vertices;
	std::vector<ModelTriangle> triangles;
	std::vector<std::vector<uint32_t>> leaf_models; // per leaf: indices into `visible`
	std::vector<std::vector<uint32_t>> leaf_decals; // per leaf: decal indices
	struct Decal
	{
		V3 pos;
		V3 axis[3]; // scaled box axes (half extents)
		uint32_t level;
		uint32_t blend;
	};
	std::vector<Decal> decals;
};

void build_dynamic_lights(swh_frame& f, const swh_map& m, const swh_frame_config& cfg);
void prepare_models(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, V3 cam_pos, const M4& view, const M4& planes);
void draw_leaf_objects(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, uint32_t leaf, const V3 clip_planes[8],
	const int32_t clip_rect[4]);

} // namespace swh
