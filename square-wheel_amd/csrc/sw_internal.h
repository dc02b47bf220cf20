// Internal records shared by the kernels and the host runtime (not part of the C ABI).
#pragma once

#include <stdint.h>

#include "../../include/sw_gpu.h"

// Screen band height of the binning / raster tiles (a tile is 64 pixels wide, one per wave).
#define BAND_H 16
// rasterizer.rs:137-271 walks at most n - 1 vertex pairs of an n <= 24 polygon.
#define MAX_PARTS_PER_DRAW 24

enum
{
	COUNTER_SPANS_LO = 0, // 64-bit span allocation counter (lo, hi)
	COUNTER_SPANS_HI = 1,
	COUNTER_POLY_PIXELS_LO = 2,
	COUNTER_POLY_PIXELS_HI = 3,
	COUNTER_TRI_PIXELS_LO = 4,
	COUNTER_TRI_PIXELS_HI = 5,
	COUNTER_BLEND_PIXELS_LO = 6,
	COUNTER_BLEND_PIXELS_HI = 7,
	COUNTER_PARTS = 8,
	COUNTER_ERROR = 9,
	COUNTER_COUNT = 16,
};

enum
{
	ERROR_BAD_SURFACE = 1u << 0,
	ERROR_SPAN_OVERFLOW = 1u << 1,
};

// A view plus the offsets of its bands / tiles / band lists inside the batch.
struct ViewExt
{
	sw_view v;
	uint32_t first_band; // global band index of band 0
	uint32_t num_bands;
	uint32_t first_tile; // global tile (wave) index of tile 0
	uint32_t cols; // 64-pixel columns
	uint64_t band_list_base; // element offset of this view's band lists (num_bands x num_draws)
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
	int32_t pad[3];
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
void launch_prep_surfaces(hipStream_t st, const sw_surface_desc* s, uint32_t n, const ViewExt* views, uint32_t nv,
	uint64_t* surf_end, uint64_t arena_texels, const MapPtrs& m, uint32_t* counters);
void launch_build_surfaces(hipStream_t st, uint32_t color_bits, const MapPtrs& m, const sw_surface_desc* s, uint32_t n,
	const uint64_t* surf_end, const ViewExt* views, uint32_t nv, const sw_light* lights, uint32_t nl, const uint32_t* lidx,
	uint32_t nli, void* arena, uint64_t arena_texels, const uint32_t* counters);
void launch_setup_draws(hipStream_t st, const sw_draw* draws, uint32_t nd, const int32_t* vdata, uint64_t nvi,
	const ViewExt* views, uint32_t nv, PartRec* parts, SpanRec* spans, DrawRec* drec, uint32_t* counters, uint64_t span_capacity,
	uint64_t arena_texels, uint64_t skybox_texels, uint64_t lite_texels);
void launch_bin_bands(hipStream_t st, const DrawRec* drec, const ViewExt* views, uint32_t nv, uint32_t total_bands,
	uint32_t* band_lists, uint32_t* band_counts, const uint32_t* counters);
void launch_raster_tiles(hipStream_t st, uint32_t color_bits, const MapPtrs& m, const sw_draw* draws, const DrawRec* drec,
	const PartRec* parts, const SpanRec* spans, const ViewExt* views, uint32_t nv, const uint32_t* band_lists,
	const uint32_t* band_counts, const void* arena, const void* skybox, void* fb, uint32_t total_tiles,
	const uint32_t* counters);
void launch_tonemap(hipStream_t st, const uint64_t* hdr, uint32_t w, uint32_t h, float isr, float isg, float isb,
	uint32_t* dst, uint32_t dst_pitch, float* sums);
} // namespace swk
