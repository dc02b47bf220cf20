// Internal records shared by the kernels and the host runtime (not part of the C ABI).
#pragma once

#include <stdint.h>

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
};

// Surface texels are built by blocks of SURF_BLOCK threads over the flattened arena.
#define SURF_BLOCK 256

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
	uint32_t stat; // DRAW_STAT_*
	uint32_t pad;
};

enum
{
	DRAW_STAT_TRIANGLE = 1u << 0,
	DRAW_STAT_READS_DST = 1u << 1, // the blending mode reads the destination
};

struct MapPtrs
{
	const sw_texel* texels;
	const uint32_t* lite;
	const float* lm_simple;
	const sw_directional_lightmap_element* lm_dir;
	const sw_texture_level* levels;
	uint32_t num_levels;
	uint32_t pad;
	uint64_t num_texels, num_lite, num_lm_simple, num_lm_dir;
};

#include <hip/hip_runtime.h>
namespace swk
{
// Kernel 1: draw setup (edge walk -> parts -> span records) with `lanes` (64 or 256) threads per
// draw and, in extra blocks, the surfaces' arena extents, validation and block -> surface table.
void launch_setup(hipStream_t st, uint32_t lanes, const sw_draw* draws, uint32_t nd, const int32_t* vdata, uint64_t nvi,
	const ViewExt* views, uint32_t nv, PartRec* parts, SpanRec* spans, DrawRec* drec, uint32_t* counters,
	uint64_t span_capacity, uint64_t arena_texels, uint64_t skybox_texels, uint64_t lite_texels,
	const sw_surface_desc* surfaces, uint32_t ns, SurfExt* surf_ext, uint32_t* blk_surf, const MapPtrs& m);
// Kernel 2: one thread per surface-arena texel.
void launch_build_surfaces(hipStream_t st, uint32_t color_bits, const MapPtrs& m, const sw_surface_desc* s, uint32_t n,
	const SurfExt* surf_ext, const uint32_t* blk_surf, const ViewExt* views, const sw_light* lights, uint32_t nl,
	const uint32_t* lidx, uint32_t nli, void* arena, uint64_t arena_texels, const uint32_t* counters);
// Kernel 3: one wave per 64 x tile_rows tile (binning fused).
void launch_raster_tiles(hipStream_t st, uint32_t color_bits, uint32_t tile_rows, const MapPtrs& m, const sw_draw* draws,
	const DrawRec* drec, const PartRec* parts, const SpanRec* spans, const ViewExt* views, uint32_t nv, const void* arena,
	const void* skybox, void* fb, uint32_t total_tiles, const uint32_t* counters);
// Kernel 4 (one block): per-draw statistics -> result block, live block cleared.
void launch_finish(hipStream_t st, const DrawRec* drec, uint32_t nd, uint32_t* counters);
void launch_tonemap(hipStream_t st, const uint64_t* hdr, uint32_t w, uint32_t h, float isr, float isg, float isb,
	uint32_t* dst, uint32_t dst_pitch, float* sums);
} // namespace swk
