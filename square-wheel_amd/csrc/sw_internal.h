// Internal records shared by the kernels and the host runtime (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sw_gpu.h"

// Raster tiles are 64 pixels wide (one pixel column per lane) and TILE_ROWS rows high, chosen per
// context among SW_TILE_ROWS_* (the autotuner times each).
#define MAX_TILE_ROWS 16
// rasterizer.rs:137-271 walks at most n - 1 vertex pairs of an n <= 24 polygon.
#define MAX_PARTS_PER_DRAW 24

// Counters. The live block holds the span allocation counter and the error flags while a batch
// runs; the finishing kernel (one block, after the raster kernel) reduces the per-draw statistics,
// writes the result block (read by the host) and zeroes the live block for the next batch, so no
// memset node is needed anywhere (also inside captured graphs).
enum
{
	COUNTER_SPANS_LO = 0, // 64-bit span allocation counter (lo, hi)
	COUNTER_SPANS_HI = 1,
	COUNTER_POLY_PIXELS_LO = 2, // result block only
	COUNTER_POLY_PIXELS_HI = 3,
	COUNTER_TRI_PIXELS_LO = 4,
	COUNTER_TRI_PIXELS_HI = 5,
	COUNTER_BLEND_PIXELS_LO = 6,
	COUNTER_BLEND_PIXELS_HI = 7,
	COUNTER_PARTS = 8,
	COUNTER_ERROR = 9,
	COUNTER_COUNT = 16, // size of one block
	COUNTER_RESULT = 16, // result block [16, 32)
	COUNTER_TOTAL = 32,
};

enum
{
	ERROR_BAD_SURFACE = 1u << 0,
	ERROR_SPAN_OVERFLOW = 1u << 1,
	ERROR_PART_GAP = 1u << 2, // the parts of a draw did not cover consecutive rows (never expected)
};

// Error words in pinned, device-mapped host memory, one per error kind (word k for the flag 1 << k):
// kernels set them with plain stores (several writers of one kind store the same value, so no
// atomic read-modify-write crosses the bus), the host reads and clears them after the sequence.
// The live counter block keeps the same flags for the later kernels of the sequence.
#define HOST_ERR_WORDS 4
#define HOST_ERR_OFFSET COUNTER_COUNT // after the result block in the host memory
#define HOST_POST_WORDS 8 // postprocess results: tonemap colour sums [0, 3), bloom average [4, 8)
#define HOST_POST_OFFSET (HOST_ERR_OFFSET + HOST_ERR_WORDS)
#define HOST_MODEL_ERR_OFFSET (HOST_POST_OFFSET + HOST_POST_WORDS) // model preparation's error word
#define HOST_WORDS_TOTAL (HOST_MODEL_ERR_OFFSET + 4)

// Surface texels are built by blocks of SURF_BLOCK threads over the flattened arena, each block
// surf_block_texels() consecutive texels: up to SURF_MAX_TEXELS_PER_THREAD per thread (the block's
// per-surface set-up is shared by more texels). A small arena (one frame) takes as few texels per
// thread as keep it within about 1024 blocks, which the device runs at once (five blocks per CU
// besides the draw blocks): a second round of blocks would add a whole block's latency.
#define SURF_BLOCK 256
#define SURF_MAX_TEXELS_PER_THREAD 8
__host__ __device__ inline uint32_t surf_block_texels(uint64_t arena_texels)
{
	const uint64_t per_thread = (arena_texels + uint64_t(SURF_BLOCK) * 1024u - 1) / (uint64_t(SURF_BLOCK) * 1024u);
	return SURF_BLOCK * uint32_t(per_thread < 1 ? 1 : per_thread > SURF_MAX_TEXELS_PER_THREAD ? SURF_MAX_TEXELS_PER_THREAD : per_thread);
}

// Arena extent of one surface and the view it belongs to.
struct SurfExt
{
	uint64_t end; // one past the surface's last arena texel
	uint32_t view;
	uint32_t pad;
};

// A view plus the offset of its tiles inside the batch.
struct ViewExt
{
	sw_view v;
	uint32_t num_bands; // tile rows of the view
	uint32_t cols; // 64-pixel columns
	uint32_t first_tile; // global tile (wave) index of tile 0
	uint32_t pad;
	uint64_t arena_end; // end of the view's arena region: the next view's offset, or the arena's end
};

// The views of a batch: a device array, or (a single view) the view itself, passed by value in
// the kernel arguments so that a one-view batch needs no upload before its launches.
struct ViewsArg
{
	const ViewExt* ptr; // device array (when !inline_one)
	uint32_t count;
	uint32_t inline_one; // the single view is `one`
	ViewExt one;
	__host__ __device__ const ViewExt& operator[](uint32_t i) const { return inline_one ? one : ptr[i]; }
};

// One part of a draw: a y range with its left/right edges (the draw_func calls of
// rasterizer.rs:202-252, the two fill_triangle_part calls of :859-965).
struct PartRec
{
	int32_t row0, row1; // rows [row0, row1) after clipping
	uint32_t span_base; // span record of row0
	uint32_t kind; // 0 perspective, 1 line z, 2 affine, 3 triangle
	int32_t d[5]; // triangle: d_tc_dx[2], d_light_dx[3] after corrections
	int32_t pad[3];
};

// One (part, row) span: clipped [x0, x1) and the interpolants at x0 with their per-pixel steps.
struct SpanRec
{
	int32_t x0, x1;
	int32_t v[6];
};

struct DrawRec
{
	uint32_t num_parts;
	int32_t row_min, row_max; // rows with non-empty spans
	int32_t x_min, x_max; // union of the non-empty spans
	uint32_t pixels; // pixels covered by the draw's spans (at most 2^28)
	uint32_t stat; // DRAW_STAT_* | kind << DRAW_STAT_KIND_SHIFT
	uint32_t rows; // span records the draw needs (written unless storage overflowed)
	uint32_t span_base; // span record of row_first; row y uses span_base + (y - row_first)
	int32_t row_first; // first row of the draw's parts
	uint32_t rows0; // rows of the first part (a triangle's lower part)
	uint32_t tex_w, tex_h; // texture size
	uint32_t pad;
	uint64_t tex_offset; // texture's first element: in the arena (view offset included), skybox or lite texels
};

// A draw's screen bounds for the raster kernel's binning (16 bytes, so that testing 64 draws
// against a tile moves little data; the full record is loaded for the draws that hit), with where
// its span records lie, so that a tile can load the span records of its draws without waiting for
// their full records. An empty draw has x_min > x_max.
struct DrawBounds
{
	int16_t x_min, x_max; // columns [x_min, x_max)
	int16_t row_min, row_max; // rows [row_min, row_max): the rows whose spans are not empty
	uint32_t span_base; // = DrawRec::span_base
	int32_t row_first; // = DrawRec::row_first
};

enum
{
	DRAW_STAT_TRIANGLE = 1u << 0,
	DRAW_STAT_READS_DST = 1u << 1, // the blending mode reads the destination
	DRAW_STAT_SKYBOX = 1u << 2, // a polygon textured from the skybox
	DRAW_STAT_PORTAL = 1u << 3, // a polygon textured from the portal textures (else the surface arena)
	DRAW_STAT_KIND_SHIFT = 4, // 2 bits: 0 perspective, 1 line z, 2 affine, 3 triangle
	DRAW_STAT_BLEND_SHIFT = 8, // 4 bits: SW_BLEND_*
};

struct MapPtrs
{
	const sw_texel* texels;
	const uint32_t* lite;
	const float* lm_simple;
	const sw_directional_lightmap_element* lm_dir;
	const sw_texture_level* levels;
	uint32_t num_levels;
	uint32_t lm_simple_offset; // byte offset of lm_simple from lm_dir (one buffer below 4 GiB)
	uint64_t num_texels, num_lite, num_lm_simple, num_lm_dir;
	const void* portal; // ColorT texels of SW_TEXSRC_PORTAL draws (set per launch)
	uint64_t num_portal;
};

// Arguments of the draw setup blocks.
struct SetupArgs
{
	const sw_draw* draws;
	uint32_t num_draws;
	uint32_t setup_blocks; // blocks of draws; the launch's later blocks do other work
	const int32_t* vdata;
	uint64_t num_vertex_ints;
	PartRec* parts;
	SpanRec* spans;
	DrawRec* drec;
	DrawBounds* bounds;
	uint32_t* counters;
	uint32_t* host_err; // HOST_ERR_WORDS error words in host memory
	uint64_t span_capacity;
	uint64_t span_stride; // > 0: draw i owns span records [i * stride, (i + 1) * stride), no allocation
	uint64_t arena_texels, skybox_texels, lite_texels, portal_texels;
};

// Arguments of the surface texel blocks.
struct SurfaceArgs
{
	const sw_surface_desc* surfaces;
	uint32_t num_surfaces;
	uint32_t num_lights;
	const SurfExt* surf_ext;
	const uint32_t* blk_surf; // block -> first surface table; null: one view, searched in the kernel
	const sw_light* lights;
	const uint32_t* light_indices;
	uint32_t num_light_indices;
	uint32_t num_shadows; // lights [0, num_shadows) have an entry in `shadows`
	void* arena;
	uint64_t arena_texels;
	uint32_t* counters; // error flags (validation of descriptors without a prepared table)
	uint32_t* host_err; // HOST_ERR_WORDS error words in host memory
	const sw_light_shadow* shadows; // validated on the host: every map lies inside shadow_texels
	const float* shadow_texels;
};

// Bloom postprocessing (postprocessor.rs:258-956): the frame's downscaled, blurred copy added to it
// before the tonemap.
#define BLOOM_MAX_RADIUS 16
#define BLOOM_KERNEL (1 + 2 * BLOOM_MAX_RADIUS)
struct BloomArgs
{
	const uint64_t* hdr; // Color64 frame, pitch = w
	uint32_t w, h;
	uint32_t bw, bh, scale_log2; // bloom buffer size, scale 2^scale_log2 (1..3)
	int32_t radius; // blur radius (0: no blur)
	uint64_t* buf0; // bloom buffer (downscaled, then blurred)
	uint64_t* buf1; // horizontal pass output
	int32_t* line_sums; // 4 per bloom row: the row's sum of averages, lanes B, G, R, A (16-byte aligned)
	float* average; // 4 floats: the average colour, lanes B, G, R, A (host memory)
	int32_t k1[BLOOM_KERNEL], k2[BLOOM_KERNEL]; // integer weights of the two passes (scale 2^14)
	float inv_scale[3]; // tonemap 1 / (exposure * modulate), lanes B, G, R
	uint32_t linear; // bilinear upsampling of the bloom buffer (else nearest)
	uint32_t* dst;
	uint32_t dst_pitch; // in pixels
};

namespace swk
{
// Events a timed stage offers to its kernel launch (profiling only): the launch records them at the
// kernel's start and end and sets `used`; per host thread.
struct LaunchEvents
{
	hipEvent_t start = nullptr, stop = nullptr;
	bool used = false;
};
LaunchEvents& launch_events();

// Kernel 1: draw setup (edge walk -> parts -> span records) with `lanes` (64 or 256) threads per
// draw and, in extra blocks, the surfaces' arena extents, validation and block -> surface table
// (device-resident descriptors).
void launch_setup(hipStream_t st, uint32_t lanes, SetupArgs a, const ViewsArg& views, const sw_surface_desc* surfaces,
	uint32_t ns, SurfExt* surf_ext, uint32_t* blk_surf, const MapPtrs& m);
// Kernel 2: one thread per surface-arena texel.
void launch_build_surfaces(hipStream_t st, uint32_t color_bits, const MapPtrs& m, const SurfaceArgs& sa, const ViewsArg& views,
	const uint32_t* counters);
// Kernels 1 + 2 in one launch when the host prepared the surfaces.
void launch_setup_and_surfaces(hipStream_t st, uint32_t lanes, uint32_t color_bits, SetupArgs a, const ViewsArg& views,
	const SurfaceArgs& sa, const MapPtrs& m);
// Kernel 3: one wave per 64 x tile_rows tile (binning fused); scalar_spans: span records through
// the scalar unit (else vector loads shared with readlane).
void launch_raster_tiles(hipStream_t st, uint32_t color_bits, uint32_t tile_rows, bool scalar_spans, const MapPtrs& m,
	const DrawBounds* bounds, const DrawRec* drec, const PartRec* parts, const SpanRec* spans, const ViewsArg& views, const void* arena,
	const void* skybox, void* fb, uint32_t total_tiles, const uint32_t* counters);
// Kernel 4 (one block): per-draw statistics -> result block (host-mapped), live block cleared.
void launch_finish(hipStream_t st, const DrawRec* drec, uint32_t nd, uint32_t* counters, uint32_t* result);
// Bloom: downscale (+ the average colour when `average`), two blur passes, tonemap with the
// upsampled bloom. Grids take a frame row per block row: height <= 65535.
void launch_bloom(hipStream_t st, const BloomArgs& a, bool average);
// Tonemap; `partials` holds 3 * TONEMAP_BLOCKS floats (per-block colour sums, every one the launch
// uses written before it is read); `totals` (3 floats, may be host memory, null: not wanted)
// receives their sum. The launch takes min(height, TONEMAP_BLOCKS) blocks.
const uint32_t TONEMAP_BLOCKS = 2048u;
void launch_tonemap(hipStream_t st, const uint64_t* hdr, uint32_t w, uint32_t h, float isr, float isg, float isb,
	uint32_t* dst, uint32_t dst_pitch, float* partials, float* totals);
// draw_mesh over prepared models: grid (triangle slots / 256, runs); bases = per run (first
// vertex int, first fan-count byte), device copies of runs and bases; the host checked every range.
void launch_expand_mesh_runs(hipStream_t st, const sw_mesh_run* runs, uint32_t num_runs, uint32_t max_triangles, const uint64_t* bases,
	const sw_model_vertex* vertices, const uint16_t* triangles, const uint32_t* counts, int32_t* out_ints, uint8_t* out_fans);
// Triangle-model preparation: one block per instance. meshes / instances / sort_offsets are device
// copies; scratch holds the sort keys of instances above model_sort_lds_limit() triangles
// (sort_offsets[i] entries into it, a power of two of them each).
void launch_prepare_models(hipStream_t st, uint32_t num_instances, const sw_model_mesh* meshes, const sw_model_instance* instances,
	const uint64_t* sort_offsets, uint64_t* scratch, sw_model_vertex* out_vertices, uint16_t* out_triangles, uint32_t* out_counts,
	uint32_t* host_err);
uint32_t model_sort_lds_limit();

// Texture operations: pack_texture_ops validates them against the level table and packs the
// launch arguments; the caller copies `packed` to the device and launches.
struct TexOpPlan
{
	size_t bytes = 0, ops_at = 0;
	uint32_t blocks = 0, num_runs = 0;
};
int pack_texture_ops(const sw_texture_op* ops, uint32_t num_ops, const sw_texture_level* levels, uint32_t num_levels,
	uint32_t num_turb_shifts, std::vector<char>& packed, TexOpPlan& plan, std::string& error);
void launch_texture_ops(hipStream_t st, const char* args_dev, const TexOpPlan& plan, const int32_t* turb_shifts_dev, sw_texel* texels,
	uint32_t* lite);
// Depth buffers: depth_plan validates and packs the draws (then targets at targets_at);
// launch_depth runs the setup (one wave per draw, row records of depth_row_record_bytes() each,
// the target's height per draw) and the raster (one wave per target row).
struct DepthPlan
{
	size_t bytes = 0, targets_at = 0;
	uint64_t row_records = 0;
	uint32_t row_blocks = 0, num_draws = 0, num_targets = 0;
};
int depth_plan(const sw_depth_target* targets, uint32_t num_targets, const sw_depth_draw* draws, uint32_t num_draws,
	uint64_t num_vertex_ints, uint64_t num_texels, std::vector<char>& packed, DepthPlan& plan, std::string& error);
size_t depth_row_record_bytes();
void launch_depth(hipStream_t st, const char* args_dev, const DepthPlan& plan, const int32_t* vdata, void* rows, int32_t* draw_rows,
	float* texels);
// A skybox side (mip 0 from the image, then up to four halvings) back to back at dst.
// water (generative_texture_effect_water.rs): one step (points, then the stencil into field_old) and
// the wavy texture + mips (dst levels given by texel offsets)
void launch_water_step(hipStream_t st, float* field, float* field_old, uint32_t log2w, uint32_t log2h, float attenuation,
	const sw_wave_point* points_dev, uint32_t num_points);
struct WaterTexLaunch
{
	uint64_t dst_offset[5]; // mip 0 .. 4
	uint64_t color_offset;
	uint32_t color_mode, init;
	float roughness;
	uint32_t base_color;
};
void launch_water_texture(hipStream_t st, const float* field, uint32_t log2w, uint32_t log2h, const WaterTexLaunch& w, sw_texel* texels);
// fire (generative_texture_effect_fire.rs): one step (update into heat_old, then the points) and the
// palette texture + lite mips (dst levels given by texel offsets)
void launch_fire_step(hipStream_t st, const uint8_t* heat, uint8_t* heat_old, uint32_t log2w, uint32_t log2h, uint32_t attenuation_i,
	bool slow, const sw_heat_point* points_dev, uint32_t num_points);
void launch_fire_texture(hipStream_t st, const uint8_t* heat, uint32_t log2w, uint32_t log2h, const uint64_t dst_offset[5],
	const uint32_t* palette_dev, uint32_t* lite);
void launch_skybox_side(hipStream_t st, uint32_t color_bits, const void* image, uint32_t image_w, uint32_t size, int32_t brightness,
	void* dst);

// One portal texture overlay, resolved against the map's level table on the host.
struct PortalMixRec
{
	uint64_t dst;        // first texel of the rect (ColorT texels)
	uint32_t w, h;       // the rect
	int32_t tcu, tcv;    // tc_start
	uint32_t tex_offset; // the lite level's first texel
	uint32_t tw, th;     // the lite level's size
	uint32_t blend;
	float light[3];      // R, G, B
	uint32_t reserved;
};
// mix_surface_with_texture over each record's rect (grid: texel chunks x records)
void launch_portal_mix(hipStream_t st, uint32_t color_bits, const uint32_t* lite, const PortalMixRec* recs, uint32_t num,
	uint64_t max_texels, void* texels);
} // namespace swk
