/*
 * sw_host.h -- C API of the CPU-side renderer front-end (libsw_host.so).
 *
 * This is the caller of the drop-in boundary in sw_gpu.h: the part of Square-Wheel's renderer that
 * stays on the CPU (visibility, per-polygon setup, draw ordering), restated in C++ because the
 * reference's Rust toolchain is not available here. Per view it emits exactly the command buffers
 * sw_gpu.h consumes:
 *
 *   swh_frame_build  <- PartialRenderer::prepare_frame + draw_frame up to the pixel loops
 *                       (partial_renderer.rs:156-261): MapVisibilityCalculator::update_visibility
 *                       (map_visibility_calculator.rs:52-97), prepare_polygon_surface (:1464-1638),
 *                       build_polygons_surfaces' per-surface parameters (:1682-1820), draw_tree /
 *                       draw_leaf / draw_polygon (:2450-2576, :4011-4142), decals and triangle models.
 *
 * It also generates the seeded content the benchmark configurations name (the reference's
 * map_compiler / lightmapper cannot run here): a voxel-grid BSP map with portals, value-noise
 * textures with normal maps and mip chains (textures.rs:93-362), and directional / simple
 * lightmaps from point lights and a sun.
 */
#ifndef SW_HOST_H
#define SW_HOST_H

#include <stdint.h>

#include "sw_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swh_map swh_map;
typedef struct swh_frame swh_frame;
typedef struct swh_shadow swh_shadow;

typedef struct swh_map_config
{
	uint32_t seed;
	uint32_t grid[3]; /* cells along x, y, z (128 units each) */
	float solid_probability;
	uint32_t num_lights; /* white point lights */
	float light_intensity_min, light_intensity_max;
	uint32_t sun; /* 1: sun shining through sky openings in the top layer */
	float sky_fraction; /* fraction of top-layer ceilings that are sky */
	uint32_t texture_size; /* 64 / 256 ... (power of two >= 16) */
	uint32_t num_materials;
	uint32_t normal_maps; /* 1: tangent-space normal maps from a Gaussian-bumped height field */
	float specular_fraction; /* share of materials with roughness U[0.1, 0.6] (half of them metal) */
	float alpha_blend_fraction; /* share of materials drawn with BlendingMode::AlphaBlend */
	uint32_t directional_lightmaps; /* 1: build DirectionalLightmapElement data too */
	uint32_t num_decal_textures; /* lite 32x32 decal textures */
	uint32_t num_model_textures; /* lite 64x64 model skins */
	uint32_t num_submodels; /* inline brush models (doors and platforms) moved by the frame time */
	uint32_t num_view_portals; /* wall portals: mirrors, camera views and parallax views, in turn */
} swh_map_config;

typedef struct swh_map_info
{
	uint32_t num_polygons, num_leafs, num_nodes, num_portals, num_empty_cells;
	uint64_t num_texels, num_lightmap_elements;
} swh_map_info;

typedef struct swh_frame_config
{
	uint32_t width, height;
	uint32_t color_bits; /* 32 or 64 */
	float fov; /* radians, horizontal (matrix.rs:31-50) */
	float textures_mip_bias;
	uint32_t use_directional_lightmaps;
	uint32_t clear_background; /* RendererConfig::clear_background */
	uint32_t force_clear; /* clear even when the camera is inside its leaf volume */
	uint32_t num_dynamic_lights; /* seeded DynamicLight (shadow None) */
	uint32_t num_models; /* seeded animated triangle models */
	uint32_t num_decals; /* seeded decals */
	uint32_t seed; /* dynamic objects seed */
	float time; /* animation time, seconds */
	uint32_t virtual_threads; /* 1: full-screen clip rect (the parity configuration) */
	uint32_t view_model; /* 1: an animated model held in front of the camera (drawn after the tree) */
	uint32_t portals_depth; /* RendererConfig::portals_depth: levels of views rendered through portals */
	/* 1: the frame's triangle models are prepared on the device. The front-end then animates no
	 * vertices and draws no model triangles: it lists the visible models as sw_model_instance
	 * (sw_prepare_models' input, the map's mesh being swh_map_model_mesh) and each mesh drawn in a
	 * leaf as an sw_mesh_run at its place in the draw list (swh_frame_model_data) */
	uint32_t device_models;
} swh_frame_config;

/* A view's command buffers, valid until the frame is rebuilt or destroyed. */
typedef struct swh_frame_data
{
	sw_view view; /* offsets zero; first_* = 0 */
	const sw_surface_desc* surfaces;
	uint32_t num_surfaces;
	const sw_light* lights;
	uint32_t num_lights;
	const uint32_t* light_indices;
	uint32_t num_light_indices;
	const sw_draw* draws;
	uint32_t num_draws;
	const int32_t* vertex_data;
	uint64_t num_vertex_ints;
	uint64_t surface_arena_texels;
	uint32_t num_visible_leafs;
	uint32_t num_triangles;
	uint32_t first_submodel_surface; /* surfaces from this index on belong to inline (brush) models */
	/* View portals (partial_renderer.rs:1834-2282): the frame's portal draws (SW_TEXSRC_PORTAL) sample
	 * num_portal_texels texels of the frame's colour type, the portal textures laid end to end. With
	 * portals_depth > 0 the textures are the framebuffers of num_portal_views views (one per visible
	 * portal, in portal order, each a frame of its own: swh_frame_portal_view), laid end to end; at
	 * the last level they are black (zero). */
	uint64_t num_portal_texels;
	uint32_t num_portal_views;
	uint32_t num_visible_portals;
} swh_frame_data;

int swh_map_create(const swh_map_config* config, swh_map** out_map);
void swh_map_destroy(swh_map* map);
/* Pointers into host arrays owned by the map (upload them with sw_map_upload). */
int swh_map_desc(const swh_map* map, sw_map_desc* out_desc);
int swh_map_info_get(const swh_map* map, swh_map_info* out_info);
/* The map's view portals (frame_info.rs ViewPortal / PortalView, placed by the game in the reference):
 * quads 1 unit in front of walls, 64 x 80 units, two texels per unit. Returns the number written
 * (at most capacity; the map's count with out == NULL). */
enum
{
	SWH_PORTAL_MIRROR = 0,
	SWH_PORTAL_CAMERA = 1, /* CameraAtPosition: a camera at camera_position (yaw, pitch, fov) */
	SWH_PORTAL_PARALLAX = 2 /* ParallaxPortal: transform maps the viewer's position to the view point */
};
typedef struct swh_view_portal
{
	uint32_t kind;
	uint32_t blending_mode;
	float plane[4]; /* vec, dist: the front side is where dot(vec, p) > dist */
	float tex_coord_equation[2][4]; /* vec, dist: tc = dot(vec, p) + dist */
	float vertices[4][3]; /* clockwise on screen seen from the front */
	float camera_position[3];
	float camera_yaw, camera_pitch, camera_fov;
	float transform[16]; /* column-major */
	/* ViewPortalTexture (frame_info.rs:198-205): texture_level is the lite level of the overlay's
	 * mip 0 (its mips follow), SW_NO_TEXTURE for none */
	uint32_t texture_level;
	uint32_t texture_blending_mode;
	float texture_light_scale;
	float texture_light_add[3];
} swh_view_portal;
int swh_map_view_portals(const swh_map* map, swh_view_portal* out, uint32_t capacity);
/* n seeded camera poses at empty-cell centres: out = n x {x, y, z, yaw, pitch}. */
int swh_camera_path(const swh_map* map, uint32_t seed, uint32_t n, float* out_poses);

int swh_frame_create(swh_frame** out_frame);
void swh_frame_destroy(swh_frame* frame);
int swh_frame_build(swh_frame* frame, const swh_map* map, const swh_frame_config* config, const float cam_pos[3],
	float yaw, float pitch);
int swh_frame_get(const swh_frame* frame, swh_frame_data* out_data);
/* A frame's portal (PortalInfo, partial_renderer.rs:1834-1996): visible when its texture has pixels;
 * the texture is resolution[0] x resolution[1] texels of the portal's texture space at `mip`, its
 * texel (0, 0) at tc_min, at texture_offset in the frame's portal textures. */
typedef struct swh_portal_info
{
	uint32_t visible;
	uint32_t resolution[2];
	int32_t tc_min[2];
	uint32_t mip;
	uint64_t texture_offset;
	/* the portal's texture overlay (partial_renderer.rs:2205-2276), for sw_mix_portal_textures:
	 * mix_level = the overlay's lite level at `mip` (SW_NO_TEXTURE: none), its blending mode and the
	 * light at the portal's centre (R, G, B) */
	uint32_t mix_level;
	uint32_t mix_blend;
	float mix_light[3];
	uint32_t reserved;
} swh_portal_info;
int swh_frame_portal_info(const swh_frame* frame, uint32_t portal_index, swh_portal_info* out_info);
/* The index-th view seen through a visible portal of the frame (index < num_portal_views); valid until
 * the frame is rebuilt or destroyed. Its own portal views follow the same scheme, one level deeper. */
int swh_frame_portal_view(const swh_frame* frame, uint32_t index, const swh_frame** out_view);

/* The frame's triangle models (prepare_models / draw_mesh, partial_renderer.rs:995-1190, 3525-3718).
 * runs[i] is a mesh drawn in one leaf; its triangles belong in the draw list just before draw
 * run_positions[i] (runs at one position keep their order). With device_models the instances are
 * sw_prepare_models' input (mesh 0 = swh_map_model_mesh, outputs at first_vertex / first_triangle,
 * counts indexed by the runs' count_index) and the draw list holds no model triangles. Without it
 * the front-end prepared the models itself: prepared_* hold its vertices, sorted triangles and
 * visible-triangle counts, the runs refer to them, and run_num_draws[i] triangle draws of the list
 * from run_positions[i] on are that run's (sw_expand_mesh_runs' fans of the same triangles). */
typedef struct swh_model_data
{
	const sw_model_instance* instances;
	uint32_t num_instances;
	const sw_mesh_run* runs;
	uint32_t num_runs;
	const uint32_t* run_positions;
	const uint32_t* run_num_draws; /* host preparation only */
	const sw_model_vertex* prepared_vertices; /* host preparation only */
	uint64_t num_prepared_vertices;
	const uint16_t* prepared_triangles; /* host preparation only, 3 per triangle */
	uint64_t num_prepared_triangles;
	const uint32_t* prepared_counts; /* host preparation only, per run count_index */
	uint32_t num_prepared_counts;
	uint64_t vertex_capacity; /* vertices and triangles the instances' outputs span */
	uint64_t triangle_capacity;
} swh_model_data;
int swh_frame_model_data(const swh_frame* frame, swh_model_data* out);
/* The draw list of a device_models frame completed with its runs' triangles: draws[0 .. num_draws)
 * with, before draw run_positions[r], run r's fan triangles (sw_expand_mesh_runs' slots with a
 * non-zero fan size, triangle t's fan in order, t in order; their vertex data at vertex int
 * int_base + the run's slot). Returns the number of draws written to out (at most capacity), or -1. */
int64_t swh_assemble_model_draws(const sw_draw* draws, uint32_t num_draws, const sw_mesh_run* runs, const uint32_t* run_positions,
	uint32_t num_runs, const uint8_t* fan_counts, uint64_t int_base, sw_draw* out, uint64_t capacity);

/* The map's animated model mesh as sw_model_mesh's VERTEX_ANIMATED arrays (host memory owned by
 * the map): tc[2] per vertex, frame_vertices = {position[3], normal[3]} per vertex per frame, and
 * triangles (3 u16 each, in the winding the camera sees front faces clockwise). */
typedef struct swh_model_mesh_data
{
	uint32_t num_vertices, num_triangles, num_frames;
	const float* tc;
	const float* frame_vertices;
	const uint16_t* triangles;
} swh_model_mesh_data;
int swh_map_model_mesh(const swh_map* map, swh_model_mesh_data* out);

/* Shadow-map draw lists (sw_render_depth's input) of the map's static geometry:
 *   swh_shadow_add_cube <- PartialRenderer::build_shadow_maps' cube branch (partial_renderer.rs:541-566)
 *                          with calculate_cube_shadow_map_side_matrices (light.rs:73-128) and
 *                          DepthRenderer::draw_map / draw_tree / draw_polygon (depth_renderer.rs:47-287):
 *                          per side, visibility from the light, back-to-front leaves, back faces
 *                          culled, the biased and scaled depth equation, projection and clipping.
 * Each call appends six size x size targets (sides +x, -x, +y, -y, +z, -z) at the running texel
 * offset; the draws keep the reference's order. Blended materials cast no shadows. */
typedef struct swh_shadow_data
{
	const sw_depth_target* targets;
	uint32_t num_targets;
	const sw_depth_draw* draws;
	uint32_t num_draws;
	const int32_t* vertex_data;
	uint64_t num_vertex_ints;
	uint64_t num_depth_texels; /* sum of the targets' width x height */
} swh_shadow_data;

int swh_shadow_create(swh_shadow** out);
void swh_shadow_destroy(swh_shadow* shadow);
void swh_shadow_clear(swh_shadow* shadow);
int swh_shadow_add_cube(swh_shadow* shadow, const swh_map* map, const float position[3], uint32_t size);
/* A projector light (PartialRenderer::build_shadow_maps' projector branch, partial_renderer.rs:567-584):
 * one size x size target with SW_DEPTH_CIRCLE, the camera from the rotation quaternion (s, x, y, z)
 * and the field of view (radians, 0 < fov < pi); basis_out (may be null) receives the sw_light_shadow
 * basis_x, basis_y, basis_z of the map (renderer_structs.rs:288-307). */
int swh_shadow_add_projector(swh_shadow* shadow, const swh_map* map, const float position[3], const float rotation[4], float fov,
	uint32_t size, float basis_out[9]);
int swh_shadow_get(const swh_shadow* shadow, swh_shadow_data* out_data);

/* Compiled-map files (common/bsp_map_save_load.rs:23-362): the map as a Square-Wheel .sqwm file,
 * format version 11 — the compact BSP map's lumps (nodes, leafs, polygons, portals, leaf portals,
 * vertices, texture names, entities with their key/value strings, lightmaps and the light grid,
 * both compressed as common/lightmap_compression.rs does). The worldspawn entity carries this
 * generator's configuration as "sw_*" keys, and the static lights are "light" entities. */
int swh_map_save_sqwm(const swh_map* map, const char* path);
/* Loads a .sqwm file written by swh_map_save_sqwm: header id and version, every lump's element
 * size and range and every index a frame follows are checked (an error, not a crash, on a bad
 * file). The textures and the voxel grid the frame builder needs are regenerated from the
 * worldspawn's generator keys; the BSP, polygons, portals, lightmaps (decompressed) and light grid
 * come from the file. */
int swh_map_load_sqwm(const char* path, swh_map** out_map);
/* The map's geometry as a Quake-format MAP file (common/map_file_q1.rs:24-115) for the reference's
 * map_compiler: one axis-aligned box brush per run of solid cells along x (sky cells with the sky
 * texture), the static lights as "light" entities, an info_player_start. */
int swh_map_save_quake_map(const swh_map* map, const char* path);
/* Message of the last failed save / load call. */
const char* swh_last_error(void);
/* CompressedColor / CompressedVector (lightmap_compression.rs), exposed for the tests:
 * out = {components[3], scale}. */
void swh_compress_color(const float color[3], uint8_t out[4]);
void swh_decompress_color(const uint8_t compressed[4], float out[3]);
void swh_compress_vector(const float v[3], int8_t out_components[3], uint8_t* out_scale);

/* Per-frame texture animation parameters (the map materials processor's host side, for
 * sw_apply_texture_ops). swh_turb_shifts: make_turb_distortion's shift table for one mip
 * (generative_texture_effect_turb.rs:98-112) in a buffer kept between calls like the reference's
 * temp buffer: it becomes max(width, height) long (entries past previous_len zero, the others
 * kept) and its first `width` entries are recomputed; returns the new length. The buffer holds at
 * least max(width, height) entries. */
uint32_t swh_turb_shifts(
	float amplitude, float wave_length, float frequency, float time_s, uint32_t width, uint32_t height, uint32_t mip, int32_t* buffer,
	uint32_t previous_len);
/* SingleArgumentFunction::evaluate (common/material_function.rs:90-156): kind 0 constant (p[0]),
 * 1 linear (scale, offset), 2 sine, 3 triangle, 4 saw-tooth, 5 step (frequency, phase, amplitude,
 * offset). */
float swh_evaluate_function(uint32_t kind, const float p[4], float x);
/* A layer's texture-coordinate shift at a mip: (value as i32) >> mip. */
int32_t swh_layer_shift(float value, uint32_t mip);

/* Map-side building blocks, exposed for the tests. */
/* make_texture (textures.rs:93-161) of a single texel: normal map RGB + roughness -> packed. */
int32_t swh_pack_normal_roughness_from_rgb(uint32_t normal_rgb, float roughness);
/* build_texture_mips (textures.rs:288-362): mip0 (w x h) -> mips 1..4 back to back. */
uint64_t swh_build_texture_mips(const sw_texel* mip0, uint32_t w, uint32_t h, sw_texel* out);

#ifdef __cplusplus
}
#endif

#endif /* SW_HOST_H */
