/*
 * sw_gpu.h -- C ABI of the MI355X surface builder, span rasteriser and HDR tonemap.
 *
 * This is the drop-in boundary for the per-frame pixel path of Square-Wheel's software renderer.
 * The caller (the renderer front-end: visibility, per-polygon setup, draw ordering) stays on the
 * CPU and hands over two command buffers per view:
 *
 *   - an array of surface descriptors  (what square_wheel/partial_renderer.rs:1682-1820 passes to
 *     build_surface_{simple,directional}_lightmap and mix_surface_with_texture), and
 *   - an ordered draw list              (what partial_renderer.rs:2311-2792 / 4011-4142 passes to
 *     Rasterizer::fill_polygon and Rasterizer::fill_triangle, rasterizer.rs:35, 731).
 *
 * All structs are plain C PODs (#[repr(C)]-compatible), all functions return 0 on success and a
 * negative sw_status on failure (the message is kept in sw_last_error), nothing throws across the
 * ABI. One context per host thread; work is queued on the context's HIP stream and completes at
 * sw_sync / sw_frame_end / sw_tonemap.
 *
 * Colour layout is the reference's (common/color.rs:2-168): Color32 = A<<24|R<<16|G<<8|B,
 * Color64 = A<<48|R<<32|G<<16|B. Framebuffer pitch is given in bytes here (the reference's
 * SurfaceInfo::pitch is in pixels, common/system_window.rs:11-16).
 */
#ifndef SW_GPU_H
#define SW_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * Status codes
 * ------------------------------------------------------------------------------------------- */
typedef enum sw_status
{
	SW_OK = 0,
	SW_ERR_INVALID_ARGUMENT = -1,
	SW_ERR_OUT_OF_MEMORY = -2,
	SW_ERR_DEVICE = -3,
	SW_ERR_STATE = -4,
	SW_ERR_BUDGET = -5,
	SW_ERR_NOT_LOADED = -6,
} sw_status;

/* BlendingMode, common/material.rs:102-116. */
enum
{
	SW_BLEND_NONE = 0,
	SW_BLEND_AVERAGE = 1,
	SW_BLEND_ADDITIVE = 2,
	SW_BLEND_ALPHA_TEST = 3,
	SW_BLEND_ALPHA_BLEND = 4,
};

/* TetureCoordinatesInterpolationMode, rasterizer.rs:5-14. */
enum
{
	SW_TC_PERSPECTIVE = 0,
	SW_TC_LINE_Z = 1,
	SW_TC_AFFINE = 2,
};

/* Draw kinds. */
enum
{
	SW_DRAW_POLYGON = 0,  /* Rasterizer::fill_polygon, texture = built surface or skybox side */
	SW_DRAW_TRIANGLE = 1, /* Rasterizer::fill_triangle, texture = TextureLite (Color32) */
};

/* Polygon texture sources. */
enum
{
	SW_TEXSRC_SURFACE = 0, /* offset into the frame's surface arena (ColorT texels) */
	SW_TEXSRC_SKYBOX = 1,  /* offset into the map's skybox texel buffer (ColorT texels) */
	SW_TEXSRC_PORTAL = 2,  /* offset into the batch's portal textures (ColorT texels): the framebuffers of
	                          an earlier batch that rendered the views seen through portals and mirrors */
};

/* Texture flags (textures.rs:11-18). */
enum
{
	SW_TEX_HAS_NORMAL_MAP = 1u << 0,
	SW_TEX_HAS_NON_ONE_ROUGHNESS = 1u << 1,
	SW_TEX_IS_METAL = 1u << 2,
	SW_TEX_LITE_LEVEL = 1u << 8, /* the level indexes lite_texels (Color32), not texels */
};

/* Surface flags. */
enum
{
	SW_SURF_DIFFUSE_LAYER = 1u << 0, /* material.light && material.diffuse (partial_renderer.rs:1701) */
	SW_SURF_DIRECTIONAL = 1u << 1,   /* use DirectionalLightmapElement data (partial_renderer.rs:1677) */
	SW_SURF_EMISSIVE = 1u << 2,      /* apply mix_surface_with_texture (partial_renderer.rs:1801-1819) */
};

#define SW_MAX_POLYGON_LIGHTS 6       /* partial_renderer.rs:1666 */
#define SW_MAX_POLYGON_VERTICES 24    /* renderer_utils.rs:4 */
#define SW_MAX_LIGHTMAP_SIZE 17       /* common/lightmap.rs:9; line buffer holds +2 */
#define SW_NO_TEXTURE 0xFFFFFFFFu

/* ---------------------------------------------------------------------------------------------
 * Map-resident data (uploaded once, kept in HBM)
 * ------------------------------------------------------------------------------------------- */

/* TextureElement, textures.rs:20-26: diffuse Color32 + PackedNormalRoughness (i8 x4, scale 127). */
typedef struct sw_texel
{
	uint32_t diffuse;
	int32_t packed_normal_roughness;
} sw_texel;

/* One mip level of a Texture (textures.rs:11-18) or of a TextureLite (textures.rs:89-91). */
typedef struct sw_texture_level
{
	uint64_t offset; /* element offset into the texel buffer */
	uint32_t width;
	uint32_t height;
	uint32_t flags; /* SW_TEX_* (unused for lite textures) */
	uint32_t reserved;
} sw_texture_level;

/* DirectionalLightmapElement, common/bsp_map_compact.rs:144-156 (40 bytes). */
typedef struct sw_directional_lightmap_element
{
	float ambient_light[3];
	float light_direction_vector_scaled[3];
	float directional_light_deviation;
	float directional_light_color[3];
} sw_directional_lightmap_element;

typedef struct sw_map_desc
{
	const sw_texel* texels; /* all Texture mip levels */
	uint64_t num_texels;
	const uint32_t* lite_texels; /* all TextureLite mip levels (Color32) */
	uint64_t num_lite_texels;
	const float* lightmaps_simple; /* LightmapElement = [f32;3], bsp_map_compact.rs:142 */
	uint64_t num_lightmap_elements;
	const sw_directional_lightmap_element* lightmaps_directional;
	uint64_t num_directional_lightmap_elements;
	const sw_texture_level* levels; /* texture table (textures and lite textures, all mips) */
	uint32_t num_levels;
	uint32_t reserved;
	const uint32_t* skybox_texels32; /* skybox side textures, Color32 (may be NULL) */
	uint64_t num_skybox_texels32;
	const uint64_t* skybox_texels64; /* skybox side textures, Color64 (may be NULL) */
	uint64_t num_skybox_texels64;
} sw_map_desc;

/* ---------------------------------------------------------------------------------------------
 * Per-frame command buffers
 * ------------------------------------------------------------------------------------------- */

/* DynamicLightWithShadow with ShadowMap::None (light.rs:4-11, renderer_structs.rs:228-264). */
typedef struct sw_light
{
	float position[3];
	float radius;
	float inv_square_radius;
	float color[3];
} sw_light;

/*
 * One visible polygon's surface, as prepared by PartialRenderer::build_polygons_surfaces
 * (partial_renderer.rs:1682-1820): basis already scaled for the mip and corrected by the
 * texture-scroll shift (:1736-1740), tc_start = surface_tc_min + shift (:1690-1693).
 */
typedef struct sw_surface_desc
{
	float basis_u[3];
	float basis_v[3];
	float basis_start[3];
	float basis_normal[3];
	uint32_t size[2];
	int32_t tc_start[2];
	uint32_t texture_level; /* index into sw_map_desc::levels (the mip level to sample) */
	uint32_t lightmap_offset; /* element offset; 0 = no lightmap (partial_renderer.rs:1757-1765) */
	uint32_t lightmap_size[2];
	uint32_t lightmap_scale_log2; /* LIGHTMAP_SCALE_LOG2 - mip, in [0, 4] */
	uint32_t lightmap_tc_shift[2];
	uint32_t first_light; /* index into the frame's light-index list */
	uint32_t num_lights; /* <= SW_MAX_POLYGON_LIGHTS */
	uint32_t emissive_level; /* lite texture level, SW_NO_TEXTURE if none */
	float emissive_light[3];
	uint32_t emissive_blend; /* SW_BLEND_* */
	uint32_t flags; /* SW_SURF_* */
	uint64_t pixels_offset; /* element offset into the surface arena (ColorT elements) */
} sw_surface_desc;

/*
 * One draw of the ordered draw list. Polygon draws are exactly the arguments of
 * Rasterizer::fill_polygon after draw_polygon's setup (partial_renderer.rs:4011-4142):
 * clipped convex clockwise vertices in Fixed16 (vertex_data: x0,y0,x1,y1,...), the (possibly
 * z-scaled) depth and texture-coordinate equations, the interpolation mode and blending mode.
 * Triangle draws are the arguments of Rasterizer::fill_triangle: three TrianglePointProjected
 * (rasterizer.rs:1491-1498), 7 ints each {x, y, tc0, tc1, light0, light1, light2}.
 */
typedef struct sw_draw
{
	uint32_t kind; /* SW_DRAW_* */
	uint32_t blend; /* SW_BLEND_* */
	uint32_t mode; /* SW_TC_* (polygons) */
	uint32_t texture_source; /* SW_TEXSRC_* (polygons) */
	uint32_t first_vertex_int; /* offset into vertex_data (ints) */
	uint32_t num_vertices; /* polygon: 3..24; triangle: 3 */
	float depth_equation[3]; /* d_inv_z_dx, d_inv_z_dy, k (equations.rs:5-10) */
	float tc_equation[6]; /* d_tc_dx[2], d_tc_dy[2], k[2] (equations.rs:14-19) */
	uint64_t texture_offset; /* element offset: surface arena / skybox buffer / lite texels */
	int32_t texture_size[2];
	int32_t clip_rect[4]; /* min_x, min_y, max_x, max_y (rasterizer.rs:1476-1482) */
} sw_draw;

/* One view (frame) of a batch. */
typedef struct sw_view
{
	uint32_t width;
	uint32_t height;
	float cam_pos[3];
	uint32_t clear_background; /* fill with background_color before drawing */
	uint64_t background_color; /* raw ColorT value */
	uint32_t first_surface, num_surfaces;
	uint32_t first_draw, num_draws;
	uint64_t fb_offset; /* element offset of this view's framebuffer in the batch output (pitch = width) */
	uint64_t surface_arena_offset; /* added to every surface pixels_offset / polygon texture_offset */
} sw_view;
/* The views' surface_arena_offset values are non-decreasing. A view's arena region runs from its
 * offset to the next view's (the last view's to the arena's end): its surfaces lie inside it (an
 * error otherwise) and its polygon draws read their surface texels inside it (a draw that does not
 * is skipped). */

/*
 * A batch of views in one call. Pointers are host pointers (inputs_on_device = 0) or device
 * pointers (inputs_on_device = 1, e.g. command buffers kept resident in HBM by the caller).
 */
typedef struct sw_batch
{
	const sw_view* views;
	uint32_t num_views;
	uint32_t color_bits; /* 32 or 64 */
	const sw_surface_desc* surfaces;
	uint32_t num_surfaces;
	uint32_t num_lights;
	const sw_light* lights;
	const uint32_t* light_indices;
	uint32_t num_light_indices;
	uint32_t num_draws;
	const sw_draw* draws;
	const int32_t* vertex_data;
	uint64_t num_vertex_ints;
	uint64_t surface_arena_texels; /* total arena size in ColorT elements */
	uint64_t framebuffer_texels; /* total output size in ColorT elements */
	uint32_t inputs_on_device;
	uint32_t reserved;
	/* Textures of SW_TEXSRC_PORTAL draws (build_view_portals_textures, partial_renderer.rs:2008-2282):
	 * a device buffer of ColorT texels, typically the output of an earlier sw_render_batch over the
	 * portal views; NULL / 0 when no draw samples it. */
	const void* portal_texels;
	uint64_t num_portal_texels;
} sw_batch;

/* Work counters of the last batch (used for the algorithmic-bytes roofline). */
typedef struct sw_stats
{
	uint64_t surface_texels; /* P_surf */
	uint64_t lightmap_elements; /* L: lightmap elements read (two rows per surface row) */
	uint64_t emissive_texels; /* P_emis */
	uint64_t polygon_pixels; /* P_poly */
	uint64_t triangle_pixels; /* P_tri */
	uint64_t blend_pixels; /* P_blend: writes whose blending mode reads the destination */
	uint64_t spans; /* span records produced */
	uint64_t parts; /* polygon/triangle parts produced by the edge walk */
	uint64_t reruns; /* raster re-runs after span storage overflowed (work queued behind the batch saw the first run) */
} sw_stats;

typedef struct sw_ctx sw_ctx;
typedef struct sw_map sw_map;

/* ---------------------------------------------------------------------------------------------
 * Context and map.  (Replaces Renderer::new, renderer.rs:24-59, and the device-side half of the
 * resources manager: textures/mips and lightmaps become HBM-resident.)
 * ------------------------------------------------------------------------------------------- */
int sw_ctx_create(int device, sw_ctx** out_ctx);
void sw_ctx_destroy(sw_ctx* ctx);
const char* sw_last_error(const sw_ctx* ctx);
int sw_sync(sw_ctx* ctx);

int sw_map_upload(sw_ctx* ctx, const sw_map_desc* desc, sw_map** out_map);
void sw_map_release(sw_ctx* ctx, sw_map* map);

/* ---------------------------------------------------------------------------------------------
 * Plain drop-in path (one view per frame).
 *   sw_frame_begin      <- Renderer::prepare_frame<ColorT>   (renderer.rs:61-104)
 *   sw_build_surfaces   <- PartialRenderer::build_polygons_surfaces (partial_renderer.rs:1640-1832)
 *   sw_raster_draws     <- PartialRenderer::perform_rasterization   (partial_renderer.rs:1192-1325)
 *   sw_frame_end        <- Renderer::draw_frame's write into &mut [ColorT] (renderer.rs:106-124)
 *   sw_tonemap          <- Postprocessor::perform_postprocessing, bloom off (postprocessor.rs:68-204)
 * sw_build_surfaces copies the descriptors and returns; the surfaces are built on the device by
 * the launch sequence of the next sw_raster_draws (or sw_frame_end / sw_read_surfaces), ahead of
 * the draws that sample them, so a frame costs three kernel launches.
 * ------------------------------------------------------------------------------------------- */
int sw_frame_begin(
	sw_ctx* ctx, sw_map* map, uint32_t color_bits, uint32_t width, uint32_t height, const float cam_pos[3],
	uint32_t clear_background, uint64_t background_color);
int sw_build_surfaces(
	sw_ctx* ctx, const sw_surface_desc* surfaces, uint32_t num_surfaces, const sw_light* lights, uint32_t num_lights,
	const uint32_t* light_indices, uint32_t num_light_indices);
/* The portal textures that SW_TEXSRC_PORTAL draws of the plain path sample (device ColorT texels;
 * NULL / 0: none), kept for the context's following frames. */
int sw_set_portal_textures(sw_ctx* ctx, const void* portal_texels_device, uint64_t num_portal_texels);

/* A portal texture's overlay: mix_surface_with_texture of a lite texture level over the portal's
 * rect of the portal textures, after its view has been drawn there and before any draw samples it
 * (the portal.texture case of build_view_portals_textures, partial_renderer.rs:2205-2276).
 * Texel (x, y) of the rect takes the level's texel ((tc_start[0] + x) rem_euclid w, (tc_start[1] + y)
 * rem_euclid h), modulated by cvtps2dq(light * 256) per channel, blended by `blend`. */
typedef struct sw_portal_mix
{
	uint64_t dst_offset; /* the rect's first texel in the portal textures; row pitch = size[0] */
	uint32_t size[2];
	int32_t tc_start[2]; /* the portal's tc_min at its mip */
	uint32_t level;      /* a lite level of the map (the portal texture at the portal's mip) */
	uint32_t blend;      /* SW_BLEND_* */
	float light[3];      /* R, G, B: light_scale x the light grid's light at the portal centre + light_add */
	uint32_t reserved;
} sw_portal_mix;

/* Applies the overlays in order (rects of one call must not overlap) to device ColorT texels. */
int sw_mix_portal_textures(sw_ctx* ctx, sw_map* map, uint32_t color_bits, const sw_portal_mix* mixes, uint32_t num_mixes,
	void* portal_texels_device, uint64_t num_portal_texels);

int sw_raster_draws(
	sw_ctx* ctx, const sw_draw* draws, uint32_t num_draws, const int32_t* vertex_data, uint64_t num_vertex_ints);
int sw_frame_end(sw_ctx* ctx, void* dst, size_t dst_pitch_bytes, uint32_t dst_on_device);
int sw_tonemap(
	sw_ctx* ctx, float exposure, const float color_modulate[3], void* dst32, size_t dst_pitch_bytes,
	uint32_t dst_on_device, float out_colors_sum[3]);

/* ---------------------------------------------------------------------------------------------
 * Batched path: many views of uneven size in one set of launches (surfaces of all views in one
 * flattened-texel launch, all draws of all views in one setup launch, one raster launch).
 * out_framebuffer is a device pointer of batch->framebuffer_texels ColorT elements.
 * No counterpart in the reference (its renderer draws one view per call); it exists for callers
 * that render many independent views (multi-view serving, cube maps, portals).
 * ------------------------------------------------------------------------------------------- */
int sw_render_batch(sw_ctx* ctx, sw_map* map, const sw_batch* batch, void* out_framebuffer_device);

/* Asynchronous form of sw_render_batch, for schedulers that keep several batches in flight on
 * several contexts: validates, uploads host-resident inputs (copied before returning) and queues
 * the launches on the context's stream, then returns. sw_batch_poll returns 1 once the batch has
 * finished on the device (0 while it runs); sw_batch_wait waits, checks the work counters (re-running
 * the raster stage if span storage overflowed) and reports the batch's status; stats are valid
 * after it. One batch in flight per context. No counterpart in the reference (one synchronous
 * frame per call). */
int sw_render_batch_async(sw_ctx* ctx, sw_map* map, const sw_batch* batch, void* out_framebuffer_device);
int sw_batch_poll(sw_ctx* ctx);
int sw_batch_wait(sw_ctx* ctx);

/* Captures the last sw_render_batch call's launches (same shapes, same device buffers) into a
 * HIP graph and replays it; inputs must be device-resident and unchanged in size. */
int sw_render_batch_graph(sw_ctx* ctx, sw_map* map, const sw_batch* batch, void* out_framebuffer_device);

/* Raster tile height. The raster kernel gives each 64-pixel-wide screen tile of 4, 8 or 16 rows
 * to one wave (default 8). Small tiles give more waves (latency hiding on small frames), tall tiles
 * amortise the per-tile draw binning (large batches). No counterpart in the reference (its
 * rasteriser walks the rows of one screen rect per thread, rect_splitting.rs:23-78). */
#define SW_TILE_ROWS_COUNT 3 /* candidates: 4, 8, 16 */
int sw_set_tile_rows(sw_ctx* ctx, uint32_t rows);
uint32_t sw_get_tile_rows(const sw_ctx* ctx);

/* How the raster kernel loads the span records of its (draw, row) items: 1 = one scalar load per
 * item into SGPRs, 0 = one vector load per 16 items shared with readlane. Both give identical
 * output; which is faster depends on the workload (the autotuner picks). */
int sw_set_scalar_spans(sw_ctx* ctx, uint32_t enable);
uint32_t sw_get_scalar_spans(const sw_ctx* ctx);

/* Autotuner: renders `batch` with every tile height and both span load modes, keeps the fastest
 * for this context and reports the milliseconds per batch of each tile height (4, 8, 16 rows; the
 * better mode) in out_ms (may be NULL). */
int sw_autotune_batch(
	sw_ctx* ctx, sw_map* map, const sw_batch* batch, void* out_framebuffer_device, uint32_t* out_rows,
	double out_ms[SW_TILE_ROWS_COUNT]);
/* The same over a sample of batches (each candidate timed over all of them, median of five
 * rounds; out_ms per batch): for a path whose frames differ, such as a camera path. out_framebuffer
 * holds the largest batch's framebuffer. */
int sw_autotune_batches(
	sw_ctx* ctx, sw_map* map, const sw_batch* batches, uint32_t num_batches, void* out_framebuffer_device,
	uint32_t* out_rows, double out_ms[SW_TILE_ROWS_COUNT]);

/* Tonemap a device Color64 buffer (pitch = width) into a device Color32 buffer. */
int sw_tonemap_device(
	sw_ctx* ctx, const void* hdr_device, uint32_t width, uint32_t height, float exposure,
	const float color_modulate[3], void* dst32_device, size_t dst_pitch_bytes, float out_colors_sum[3]);

/* No counterpart in the reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). sw_tonemap_device with the same output bits,
 * allocating nothing: the per-block partial colour sums go to a static device array of
 * 3 * 2048 floats (24 KiB per GPU, held until the process ends) instead of a context buffer, and the
 * three colour sums (may be NULL) to a device pointer, without a host synchronisation; the work is
 * queued on the context's stream. Every call on one GPU shares that array: calls that overlap on one
 * GPU (on several contexts or streams) race. Partial sums beyond the array's capacity are rejected
 * (the launch never needs more than 2048 blocks' worth). */
int sw_tonemap_device_noalloc(
	sw_ctx* ctx, const void* hdr_device, uint32_t width, uint32_t height, float exposure,
	const float color_modulate[3], void* dst32_device, size_t dst_pitch_bytes, float* out_colors_sum_device);

/* The bloom settings of PostprocessorConfig (postprocessor_config.rs), as the reference clamps them
 * (postprocessor.rs:928-940): sigma in [0, 40] (0 = no bloom), scale_log2 in [1, 3], scale in
 * [1/16, 1/2]. */
typedef struct sw_bloom_params
{
	float sigma; /* bloom_sigma, in frame pixels */
	uint32_t scale_log2; /* bloom_buffer_scale_log2 */
	float scale; /* bloom_scale: the blurred buffer's weight */
	uint32_t linear_filter; /* linear_bloom_filter: bilinear upsampling (else nearest) */
} sw_bloom_params;

/* Postprocessor::perform_postprocessing (postprocessor.rs:68-204) on a device Color64 frame:
 * with bloom (bloom != NULL and sigma > 0): the frame is downscaled by 2^scale_log2 (integer box
 * average), blurred by a separable integer Gaussian and added, upsampled, to every pixel before the
 * tonemap (:258-956); without, the plain tonemap. out_average_rgb: the average colour the reference
 * hands to update_exposure (the downscaled buffer's with bloom, the frame's without). The frame
 * must be at least one bloom texel in each direction and at most 65535 rows tall; dst_pitch_bytes
 * a multiple of 4, at least width * 4. Synchronous when out_average_rgb is given (it waits for the
 * average), else queued on the context's stream. */
int sw_postprocess_device(
	sw_ctx* ctx, const void* hdr_device, uint32_t width, uint32_t height, float exposure, const float color_modulate[3],
	const sw_bloom_params* bloom, void* dst32_device, size_t dst_pitch_bytes, float out_average_rgb[3]);

/* The auto-exposure settings of PostprocessorConfig, clamped as the reference does. */
typedef struct sw_exposure_params
{
	float exposure_update_speed;
	float base_brightness;
	float zero_level_brightness;
	float brightness_scale_power;
	float min_exposure;
	float max_exposure;
} sw_exposure_params;

/* ---------------------------------------------------------------------------------------------
 * Shadow maps of dynamic lights (light.rs:4-60, 185-294)
 * ------------------------------------------------------------------------------------------- */
#define SW_SHADOW_NONE 0u
#define SW_SHADOW_CUBE 1u /* six size x size sides back to back, in CubeMapSide order (+x, -x, +y, -y, +z, -z) */
#define SW_SHADOW_PROJECTOR 2u /* one size x size map seen along basis_z */

/* A light's ShadowMap: depth values (1 / z, larger is nearer) in a device float buffer. */
typedef struct sw_light_shadow
{
	uint32_t kind; /* SW_SHADOW_* */
	uint32_t size;
	uint64_t offset; /* element offset of the map (cube: of side 0) in the shadow texel buffer */
	float basis_x[3], basis_y[3], basis_z[3]; /* projector */
	uint32_t reserved;
} sw_light_shadow;

/* The shadow maps the surface builder samples from the next launch on (until called again; num 0
 * turns them off): shadows[i] belongs to light i of each frame's light list (lights past num cast
 * no shadow); shadow_texels_device holds num_shadow_texels floats (e.g. filled by sw_render_depth).
 * get_light_shadow_factor (light.rs:185-294) becomes a 0/1 factor of the light's distance term. */
int sw_set_light_shadows(
	sw_ctx* ctx, const sw_light_shadow* shadows, uint32_t num, const float* shadow_texels_device, uint64_t num_shadow_texels);

/* A polygon of a depth buffer (DepthRenderer::draw_polygon, depth_renderer.rs:199-287, drawn by
 * DepthRasterizer::fill_polygon, rasterizer.rs:1295-1474): convex clockwise Fixed16 vertices
 * (x, y pairs in the vertex data) and the DepthEquation of 1 / z. */
typedef struct sw_depth_draw
{
	uint32_t first_vertex_int; /* index of the first vertex's x in the vertex data */
	uint32_t num_vertices; /* 3 .. SW_MAX_POLYGON_VERTICES */
	float depth_equation[3]; /* d_inv_z_dx, d_inv_z_dy, k */
	uint32_t depth_test; /* 1: keep the larger 1 / z (submodels); 0: overwrite (back-to-front order) */
} sw_depth_draw;

/* A depth buffer (a cube side or a projector map) and its draws, which follow the previous
 * target's in the draw array. */
typedef struct sw_depth_target
{
	uint64_t offset; /* of the width x height buffer in the depth texels */
	uint32_t width, height; /* at most 65535 */
	uint32_t first_draw, num_draws;
	uint32_t flags; /* SW_DEPTH_* */
	uint32_t reserved;
} sw_depth_target;

/* After the target's draws, make_shadow_map_circle (light.rs:130-154): the texels outside the
 * square's inscribed circle become 1e24 (a projector's round beam). Square targets only. */
#define SW_DEPTH_CIRCLE 1u

/* Draws every target's polygons in order into its buffer (no clear: pixels no polygon covers keep
 * their values, as the reference's reused shadow-map storage does). draws, vertex_data and targets
 * are host arrays; depth_texels_device holds num_depth_texels floats (the shadow texels of
 * sw_set_light_shadows). Queued on the context's stream. */
int sw_render_depth(sw_ctx* ctx, const sw_depth_target* targets, uint32_t num_targets, const sw_depth_draw* draws, uint32_t num_draws,
	const int32_t* vertex_data, uint64_t num_vertex_ints, float* depth_texels_device, uint64_t num_depth_texels);

/* ---------------------------------------------------------------------------------------------
 * Per-frame texture animation (the map materials processor's generative effects) on the map's
 * resident textures: the caller (the materials processor) evaluates the effect's time functions
 * and hands over one operation per mip level.
 * ------------------------------------------------------------------------------------------- */
#define SW_TEXOP_TURB 0u /* make_turb_distortion (generative_texture_effect_turb.rs:89-138) */
#define SW_TEXOP_LAYER 1u /* apply_texture_layer (generative_texture_effect_layered.rs:162-381) */
#define SW_TEXOP_LAYER_LITE 2u /* an emissive layer: mix_surface_with_texture (surfaces.rs:854-971) */

typedef struct sw_texture_op
{
	uint32_t kind; /* SW_TEXOP_* */
	uint32_t dst_level; /* the level written (index into the map's level table) */
	uint32_t src_level; /* the level read; never a destination in the same call */
	uint32_t blend; /* layers: the layer material's SW_BLEND_* */
	int32_t shift[2]; /* layers: the texture-coordinate shift at this mip ((f(t) as i32) >> mip) */
	float light[3]; /* layers: the modulation colour (1, 1, 1: none) */
	uint32_t turb_shifts; /* TURB: index of the level's first value in turb_shifts; max(w, h) values */
} sw_texture_op;

/* Applies the operations to the map's device textures in order: operations with the same
 * destination must be consecutive (they are applied texel by texel, layer after layer); TURB and
 * TURB destinations have the source's power-of-two size, Texture levels for SW_TEXOP_LAYER,
 * TextureLite levels for SW_TEXOP_LAYER_LITE (TURB takes either, both alike). turb_shifts: the
 * per-coordinate shifts make_turb_distortion precomputes (round(sin(x * 2pi / (wave_length / 2^mip)
 * + t * frequency * 2pi) * amplitude / 2^mip) for x < width; the values past the width are the
 * reference's leftovers, see swh_turb_shifts). Queued on the context's stream. */
int sw_apply_texture_ops(
	sw_ctx* ctx, sw_map* map, const sw_texture_op* ops, uint32_t num_ops, const int32_t* turb_shifts, uint32_t num_turb_shifts);

/* make_skybox_side_texture{,64} + make_skybox_side_texture_mips (textures.rs:183-286): a skybox
 * side from a device image (Image: Color32 pixels for color_bits 32; Image64: Color64 for 64; row
 * pitch image_width) — the top-left min(w, h) square scaled by trunc(brightness * 256) / 256 per
 * channel (alpha 0), then its mips (size / 2 ... while at least 1, five levels at most) — written
 * back to back into the map's skybox buffer of that colour size from element dst_offset.
 * out_size: the side's size. Queued on the context's stream. */
int sw_make_skybox_side(sw_ctx* ctx, sw_map* map, uint32_t color_bits, const void* image_device, uint32_t image_width,
	uint32_t image_height, float brightness, uint64_t dst_offset, uint32_t* out_size);

/* Copies map texels back to host memory (waits): which = 0 the Texture texels (sw_texel), 1 the
 * TextureLite texels (Color32), 2 the Color32 skybox texels, 3 the Color64 ones; `count` elements
 * from element `first`. */
int sw_map_read(sw_ctx* ctx, const sw_map* map, uint32_t which, uint64_t first, uint64_t count, void* dst);

/* Postprocessor::update_exposure (postprocessor.rs:206-223): the next frame's exposure from this
 * frame's average colour (sw_postprocess_device) and duration. Host-side arithmetic. */
float sw_update_exposure(
	const sw_exposure_params* params, float current_exposure, const float average_rgb[3], float frame_duration_s);

/* ---------------------------------------------------------------------------------------------
 * Triangle-model preparation (triangle_models_rendering.rs:7-200, 324-391, 702-731): per model
 * instance, the mesh's vertices animated, transformed into camera space and lit, then its
 * front-facing triangles sorted back to front, for a batch of instances in one launch.
 * ------------------------------------------------------------------------------------------- */

/* VertexData's three kinds (triangle_model.rs:63-75). */
#define SW_MODEL_NON_ANIMATED 0u
#define SW_MODEL_VERTEX_ANIMATED 1u
#define SW_MODEL_SKELETON_ANIMATED 2u
#define SW_MODEL_MAX_BONES 255u /* MAX_TRIANGLE_MODEL_BONES */
#define SW_MODEL_MAX_TRIANGLES 65536u

/* A TriangleModelMesh and the parts of its TriangleModel the preparation reads, in device memory
 * (the reference's in-memory layouts):
 *   NON_ANIMATED:      vertices = VertexNonAnimated[num_vertices] {position[3], normal[3], tex_coord[2]}
 *   VERTEX_ANIMATED:   vertices = VertexAnimatedVertexConstant[num_vertices] {tex_coord[2]},
 *                      frame_vertices = VertexAnimatedVertexVariable[num_frames * num_vertices]
 *                      {position[3], normal[3]}
 *   SKELETON_ANIMATED: vertices = SkeletonAnimatedVertex[num_vertices] {tex_coord[2], position[3],
 *                      normal[3], {bone_index u8, weight u8}[4]} (40 bytes), bone_parents[num_bones]
 *                      (an index >= num_bones: no parent; a parent precedes its children),
 *                      frame_bones = column-major 4x4 matrices [num_frames * num_bones] (num_bones
 *                      0: the vertices are used as they are)
 * triangles: 3 u16 vertex indices each (Triangle, triangle_model.rs:59-61). */
typedef struct sw_model_mesh
{
	uint32_t kind;
	uint32_t num_vertices;
	uint32_t num_triangles; /* <= SW_MODEL_MAX_TRIANGLES */
	uint32_t num_frames;
	uint32_t num_bones; /* <= SW_MODEL_MAX_BONES */
	uint32_t reserved;
	const void* vertices;
	const float* frame_vertices;
	const uint16_t* triangles;
	const uint32_t* bone_parents;
	const float* frame_bones;
} sw_model_mesh;

/* One model of the frame: the arguments of animate_and_transform_triangle_mesh_vertices
 * (triangle_models_rendering.rs:7-17) and where its results go. */
typedef struct sw_model_instance
{
	uint32_t mesh; /* index into the meshes array */
	uint32_t frames[2]; /* AnimationPoint::frames (< num_frames of the mesh when it is animated) */
	float lerp; /* AnimationPoint::lerp: the weight of frames[0] */
	float model_matrix[16]; /* column-major */
	float model_view_matrix[16]; /* column-major: camera-space projection times model matrix */
	float tc_scale[2];
	float tc_shift[2];
	float light_cube[6][3]; /* ModelLightData (:398-460): the summed light cube */
	float light_vector_scaled[2][3]; /* its two directional components: static, dynamic */
	float light_color[2][3];
	uint32_t first_vertex; /* output: num_vertices sw_model_vertex from here */
	uint32_t first_triangle; /* output: up to num_triangles triangles from here */
} sw_model_instance;

/* ModelVertex3d (clipping.rs:196-201). */
typedef struct sw_model_vertex
{
	float pos[3];
	float tc[2];
	float light[3];
} sw_model_vertex;

/* animate_and_transform_triangle_mesh_vertices, then reject_triangle_model_back_faces and
 * sort_model_triangles (back to front by the largest camera-space z of the three vertices; equal
 * depths keep the mesh's triangle order — the reference's unstable sort leaves their order
 * open), for every instance. meshes and instances are host arrays (the meshes' data pointers are
 * device pointers); out_vertices (16-byte aligned, vertex_capacity entries), out_triangles (3 u16
 * per triangle, triangle_capacity triangles) and out_counts (visible triangles per instance) are
 * device memory. Queued on the context's stream; sw_sync reports an out-of-range vertex, bone or
 * parent index found on the device (the kernel clamps it). */
int sw_prepare_models(
	sw_ctx* ctx, const sw_model_mesh* meshes, uint32_t num_meshes, const sw_model_instance* instances, uint32_t num_instances,
	sw_model_vertex* out_vertices, uint64_t vertex_capacity, uint16_t* out_triangles, uint64_t triangle_capacity,
	uint32_t* out_counts);

/* One mesh drawn in one leaf from prepared models (draw_mesh, partial_renderer.rs:3525-3718): the
 * caller keeps the per-mesh decisions (the mesh box against the leaf's bounds, which 3D planes —
 * the near plane and the leaf's planes — and which 2D box planes cut it, its texture mip and
 * blending) and the run's place in the draw list; each prepared triangle is clipped by the 3D
 * planes, projected, clipped by the 2D planes and drawn as a fan of triangles. */
#define SW_MESH_MAX_PLANES3D 7u
typedef struct sw_mesh_run
{
	uint32_t count_index;    /* the instance whose out_counts entry (visible triangles) the run draws */
	uint32_t first_vertex;   /* its prepared vertices */
	uint32_t num_vertices;   /* the mesh's vertices (triangle indices are below it) */
	uint32_t first_triangle; /* its prepared, back-to-front sorted triangles (3 u16 each) */
	uint32_t max_triangles;  /* the mesh's triangle count: the run's triangle slots */
	uint32_t num_planes3d;   /* <= SW_MESH_MAX_PLANES3D; keeps dot(vec, pos) - dist >= 0 */
	uint32_t num_planes2d;   /* <= 4; keeps x * l[0] + y * l[1] - l[2] >= 0 on the projected plane */
	uint32_t blend;
	float planes3d[SW_MESH_MAX_PLANES3D][4];
	float planes2d[4][3];
	uint64_t texture_offset; /* the model's lite texture level at its mip */
	int32_t texture_size[2];
	int32_t clip_rect[4];
} sw_mesh_run;

/* The fans of the runs' triangles as triangle-draw vertex data: slot j of triangle t of run r (t below
 * max_triangles, j below fan_r = 1 + num_planes3d + num_planes2d) is 21 i32 (three vertices of
 * fixed16 x, y, tc[2], light[3]) at out_vertex_ints_device + int_base_r + 21 * (t * fan_r + j), where
 * int_base_r sums 21 * max_triangles * fan of the runs before r; out_fan_counts (host, one byte per
 * triangle slot, runs back to back) receives each triangle's fan size (0: clipped away, or beyond the
 * instance's visible triangles). The caller then draws, in its draw list at the run's place, slot
 * (t, j) for j below the fan size, t in order. Inputs are sw_prepare_models' device outputs;
 * synchronous (returns with the fan sizes in host memory). */
int sw_expand_mesh_runs(sw_ctx* ctx, const sw_mesh_run* runs, uint32_t num_runs, const sw_model_vertex* prepared_vertices_device,
	uint64_t vertex_capacity, const uint16_t* prepared_triangles_device, uint64_t triangle_capacity,
	const uint32_t* counts_device, uint32_t num_counts, int32_t* out_vertex_ints_device, uint64_t vertex_int_capacity,
	uint8_t* out_fan_counts);

/* Work counters of the last batch. The pixel counters of a batch are summed on the first call
 * after it (one small kernel and a wait), so batches that nobody asks about pay nothing for them.
 * No counterpart in the reference (its renderer keeps no such counters). */
int sw_get_stats(sw_ctx* ctx, sw_stats* out_stats);

/* Raster re-runs of the last batch after its span storage overflowed (sw_stats::reruns, without
 * summing the counters): work queued behind the batch before the wait saw the first run. No
 * counterpart in the reference. */
uint32_t sw_last_reruns(const sw_ctx* ctx);

/* Copies the first num_texels ColorT elements of the surface arena built by the last
 * sw_build_surfaces / sw_render_batch into host memory (synchronous). The arena is the renderer's
 * surfaces_pixels buffer (partial_renderer.rs:26-30); callers use it to inspect or cache built
 * surfaces. */
int sw_read_surfaces(sw_ctx* ctx, void* dst, uint64_t num_texels);

/* Device pointer of the plain path's current framebuffer (valid after sw_raster_draws). */
void* sw_frame_device_framebuffer(sw_ctx* ctx);

/* Device memory helpers for callers that keep command buffers resident in HBM (all on the
 * context's device; copies are synchronous; the copy direction follows the pointers, so a
 * device-to-device copy works through either function). */
int sw_device_alloc(sw_ctx* ctx, uint64_t bytes, void** out_device_ptr);
void sw_device_free(sw_ctx* ctx, void* device_ptr);
int sw_copy_to_device(sw_ctx* ctx, void* device_dst, const void* host_src, uint64_t bytes);
int sw_copy_from_device(sw_ctx* ctx, void* host_dst, const void* device_src, uint64_t bytes);
/* Queues a copy on the context's stream and returns (complete at the next wait / sync; truly
 * asynchronous for pinned host memory from sw_host_alloc). */
int sw_copy_from_device_async(sw_ctx* ctx, void* host_dst, const void* device_src, uint64_t bytes);
/* The other direction: the host source must stay valid and unchanged until the next wait / sync. */
int sw_copy_to_device_async(sw_ctx* ctx, void* device_dst, const void* host_src, uint64_t bytes);
/* Pinned host memory (for asynchronous copies). */
int sw_host_alloc(sw_ctx* ctx, uint64_t bytes, void** out_host_ptr);
void sw_host_free(sw_ctx* ctx, void* host_ptr);

/* Kernel timing with HIP events recorded on the context's stream around every launch.
 * Stages: 0 setup_draws (edge walk, span records and, for device-resident descriptors, surface
 * extents), 1 build_surfaces, 2 raster_tiles, 3 tonemap, 4 setup_and_surfaces (draw setup and
 * surface texels in one launch, for host-resident descriptors), 5 bloom (downscale, blur and the
 * tonemap with bloom), 6 models (sw_prepare_models), 7 textures (sw_apply_texture_ops), 8 depth
 * (sw_render_depth). sw_profile_read returns the summed milliseconds and launch counts since the
 * last reset (and resets when reset != 0). Not recorded inside captured graphs. */
/* ---------------------------------------------------------------------------------------------
 * Water: the wave-field simulation and the wavy texture of GenerativeTextureEffectWater
 * (generative_texture_effect_water.rs). The field (float32, width x height, both powers of two,
 * at least 16, area below 2^22) lives in two device buffers owned by the caller. The wave sources
 * (spots, lines, droplets and the random rain, :57-226) stay on the caller's CPU, which hands each
 * step's points over.
 * ------------------------------------------------------------------------------------------- */

/* One add_point_value of a step (:68-76): field[x, y] += value, old field[x, y] -= value, with x
 * and y wrapped (masked) to the field. */
typedef struct sw_wave_point
{
	uint32_t x, y;
	float value;
} sw_wave_point;

/* step_update_wave_field (:57-226) after its sources: the points in order, then update_wave_field
 * (:307-418) on the torus, new = ((left + right + up + down) * 0.5 - old) * attenuation, written over
 * field_old_device, as the reference writes its old field before swapping the two: the caller swaps
 * its two pointers after each call. Queued on the context's stream. */
int sw_water_step(sw_ctx* ctx, float* field_device, float* field_old_device, uint32_t log2_width, uint32_t log2_height,
	float attenuation, const sw_wave_point* points, uint32_t num_points);

#define SW_WATER_COLOR_KEEP 0u /* SingleColor / SourceTexture: the texels' colours are kept */
#define SW_WATER_COLOR_DEFORMED 1u /* SourceTextureNormalDeformed */
#define SW_WATER_COLOR_DEFORMED_X 2u /* SourceTextureNormalDeformedX */

#define SW_WATER_INIT_NONE 0u
#define SW_WATER_INIT_BASE_COLOR 1u /* the first update of SingleColor: every diffuse = base_color */
#define SW_WATER_INIT_SOURCE 2u /* the first update of SourceTexture: diffuse from color_level */

typedef struct sw_water_texture_params
{
	uint32_t dst_level; /* mip 0 (the field's size, SW_TEX_HAS_NORMAL_MAP); the next 4 levels get its mips */
	uint32_t color_level; /* the colour image (same size) for DEFORMED, DEFORMED_X and INIT_SOURCE */
	uint32_t color_mode; /* SW_WATER_COLOR_* */
	uint32_t init; /* SW_WATER_INIT_* */
	float roughness; /* the material's; clamped to [1/126, 1] as the reference does */
	uint32_t base_color; /* INIT_BASE_COLOR's colour: the source texture's last-mip texel */
} sw_water_texture_params;

/* make_wavy_texture (:420-682) from the field into the map's dst level (normal = normalize(right -
 * left, down - up, 1), packed with the roughness; colours per color_mode with the deformation scale
 * 16), then build_texture_mip (textures.rs:309-362) four times into the following levels. */
/* ---------------------------------------------------------------------------------------------
 * Fire: the heat map and the emissive texture of GenerativeTextureEffectFire
 * (generative_texture_effect_fire.rs). The heat map (bytes, width x height, powers of two, at least
 * 16, area below 2^22) lives in two device buffers owned by the caller. Heat sources, lines,
 * lightnings and particles (:96-326, random-driven) stay on the caller's CPU, which hands each
 * step's set_heat points over.
 * ------------------------------------------------------------------------------------------- */

/* One set_heat of a step (:80-86): heat[x, y] = max(heat[x, y], heat), x and y wrapped. */
typedef struct sw_heat_point
{
	uint32_t x, y;
	uint32_t heat; /* 0 .. 255 */
} sw_heat_point;

/* step_update_heat_map (:62-326): update_heat_map (:357-468; slow != 0: the row-serial variant)
 * from heat_device into heat_old_device, the last row cleared, then the points; as the reference's
 * map is updated in place, the caller swaps the two pointers after each call. Queued on the
 * context's stream. */
int sw_fire_step(sw_ctx* ctx, uint8_t* heat_device, uint8_t* heat_old_device, uint32_t log2_width, uint32_t log2_height,
	float attenuation, uint32_t slow, const sw_heat_point* points, uint32_t num_points);

/* generate_emissive_texture_based_on_heat_map (:470-486) into the map's TextureLite level dst_level
 * (the heat map's size) through the 256-colour palette, then build_texture_lite_mip
 * (textures.rs:385-415) into the following four levels. */
int sw_fire_make_texture(sw_ctx* ctx, sw_map* map, const uint8_t* heat_device, uint32_t log2_width, uint32_t log2_height,
	uint32_t dst_level, const uint32_t palette[256]);

int sw_water_make_texture(sw_ctx* ctx, sw_map* map, const float* field_device, uint32_t log2_width, uint32_t log2_height,
	const sw_water_texture_params* params);


#define SW_PROFILE_STAGES 9
int sw_profile_enable(sw_ctx* ctx, uint32_t enable);
int sw_profile_read(sw_ctx* ctx, double out_ms[SW_PROFILE_STAGES], uint64_t out_counts[SW_PROFILE_STAGES], uint32_t reset);

/* Raw HIP stream handle (hipStream_t) the context queues its work on. */
void* sw_ctx_stream(sw_ctx* ctx);

/* Bytes of HBM currently held by the context (arenas, command buffers, maps). */
uint64_t sw_ctx_device_bytes(const sw_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SW_GPU_H */
