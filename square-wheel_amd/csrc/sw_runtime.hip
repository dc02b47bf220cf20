// Host runtime of the C ABI declared in include/sw_gpu.h.
//
// Device memory model: the map (textures with their mip chains, lite textures, both lightmap
// kinds, skybox sides) is uploaded once and stays in HBM. Per batch the context owns grow-only
// device buffers for the command buffers, the surface arena and the span/part records, so
// steady-state frames allocate nothing. A batch is three launches on one stream: draw setup
// (plus surface extents), surface building, tiled raster.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "sw_internal.h"

// Scratch memory of sw_tonemap_device_noalloc: the tonemap's per-block partial colour sums (three
// floats for each of at most TONEMAP_BLOCKS blocks, 24 KiB), a static device array that every GPU
// the library runs on holds until the process ends. In host code `scratch` names a host-side stand-in,
// not a device address: its address on each device comes from hipGetSymbolAddress (scratch_address),
// and only that address is handed to kernels.
static __device__ __align__(256) float scratch[3 * swk::TONEMAP_BLOCKS];

namespace
{

// Clears the live counter block after a launch sequence without a raster kernel (which otherwise
// hands the counters over and clears them itself).
__global__ void k_reset_counters(uint32_t* counters)
{
	if(threadIdx.x < COUNTER_COUNT)
		counters[threadIdx.x] = 0;
}

__global__ void k_fill_color(void* fb, uint64_t n, uint64_t value, uint32_t bytes)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	if(bytes == 8)
		static_cast<uint64_t*>(fb)[i] = value;
	else
		static_cast<uint32_t*>(fb)[i] = uint32_t(value);
}

struct DevBuf
{
	void* p = nullptr;
	size_t cap = 0;
};

} // namespace

struct sw_map
{
	DevBuf texels, lite, lm_simple, lm_dir, levels, sky32, sky64;
	MapPtrs ptrs;
	uint64_t num_sky32 = 0, num_sky64 = 0;
	std::vector<sw_texture_level> host_levels;
};

#define ARG_STAGING_SLOTS 4u

struct sw_ctx
{
	int device = 0;
	hipStream_t stream = nullptr;
	std::string error;
	uint64_t device_bytes = 0;
	uint64_t alloc_epoch = 0; // bumped whenever a device buffer moves (invalidates captured graphs)
	uint32_t last_color_bits = 32; // colour size of the arena's current contents

	// command buffers / work buffers
	DevBuf views, surfaces, lights, light_indices, draws, vdata, surf_ext, blk_surf, parts, spans, drec, dbounds, counters, arena, fb,
		tonemap_out, tonemap_sums, bloom0, bloom1, bloom_lines, model_args, model_scratch, texop_args, shadows, depth_args, depth_vdata,
		depth_rows, depth_draw_rows, water_points, fire_args, portal_mix, mesh_args, mesh_fans;
	uint32_t num_shadows = 0; // sw_set_light_shadows: entries in `shadows`
	const float* shadow_texels = nullptr;
	uint32_t tile_rows = 8; // raster tile height (SW_TILE_ROWS_*), set by sw_set_tile_rows / the autotuner
	bool scalar_spans = true; // raster span records through the scalar unit (chosen by the autotuner)
	uint32_t* counters_host = nullptr; // pinned, device-mapped: result block (k_finish), then the error words
	float* noalloc_scratch = nullptr; // this device's address of `scratch`
	const void* portal_texels = nullptr; // sw_set_portal_textures (plain path)
	uint64_t num_portal_texels = 0;
	bool finish_launched = false; // the last launch sequence ended with k_finish (span allocation mode)
	bool stats_pending = false; // the last raster batch's statistics are computed when asked for
	uint32_t stats_draws = 0; // draws of that batch (its DrawRecs stay in `drec` until the next raster)
	uint32_t* counters_host_dev = nullptr; // its device address
	ViewsArg views_arg{}; // views of the batch being launched
	bool surface_error = false; // an invalid surface seen by a surface-only launch sequence
	bool batch_in_flight = false; // sw_render_batch_async launched, sw_batch_wait not yet called
	uint64_t span_capacity = 1u << 20;

	// plain-path frame state
	sw_map* frame_map = nullptr;
	uint32_t frame_color_bits = 32;
	sw_view frame_view{};
	uint64_t frame_arena_texels = 0;
	bool frame_active = false;
	bool frame_rastered = false;
	bool frame_surfaces_pending = false; // surfaces are built by the launch sequence of sw_raster_draws
	uint32_t frame_num_surfaces = 0, frame_num_lights = 0, frame_num_light_indices = 0;
	uint64_t frame_num_vertex_ints = 0;

	// last batch launch description (for re-runs after span overflow and graph capture)
	struct Launch
	{
		sw_map* map = nullptr;
		uint32_t color_bits = 32;
		uint32_t num_views = 0;
		const sw_surface_desc* surfaces = nullptr;
		uint32_t num_surfaces = 0;
		const sw_light* lights = nullptr;
		uint32_t num_lights = 0;
		const uint32_t* light_indices = nullptr;
		uint32_t num_light_indices = 0;
		const sw_draw* draws = nullptr;
		uint32_t num_draws = 0;
		const int32_t* vdata = nullptr;
		uint64_t num_vertex_ints = 0;
		uint64_t arena_texels = 0;
		void* arena = nullptr;
		void* fb = nullptr;
		uint32_t total_tiles = 0;
		uint32_t max_view_height = 0;
		const void* portal = nullptr; // textures of SW_TEXSRC_PORTAL draws
		uint64_t num_portal = 0;
		// how the surface kernel finds each texel's surface: SURF_TABLE_DEVICE (the setup launch
		// prepares the table), SURF_TABLE_HOST (the host did), SURF_SEARCH (one view, searched in
		// the surface blocks, which then run in the setup launch)
		uint32_t surface_mode = 0;
	} last;
	// pinned: packed launch arguments (model preparation, texture ops, depth draws, shadows), a ring
	// of slots so that the host fills the next slot while earlier copies wait in the stream
	char* arg_staging[ARG_STAGING_SLOTS] = {};
	size_t arg_staging_cap[ARG_STAGING_SLOTS] = {};
	hipEvent_t arg_staging_event[ARG_STAGING_SLOTS] = {}; // recorded after the last copy out of the slot
	uint32_t arg_staging_slot = 0; // the slot of the current stage_begin
	std::vector<char> texop_packed;
	std::vector<char> depth_packed;
	std::vector<SurfExt> host_ext;
	std::vector<uint32_t> host_blk;

	// graph replay
	hipGraphExec_t graph_exec = nullptr;
	hipGraph_t graph = nullptr;
	uint64_t graph_key = 0;
	uint64_t graph_epoch = 0;
	uint64_t graph_views_hash = 0; // content of the views array the captured graph last read

	sw_stats stats{};

	// kernel timing
	bool profiling = false;
	bool capturing = false;
	struct Pending
	{
		int stage;
		hipEvent_t a, b;
	};
	std::vector<Pending> pending;
	std::vector<hipEvent_t> event_pool;
	double prof_ms[SW_PROFILE_STAGES] = {};
	uint64_t prof_count[SW_PROFILE_STAGES] = {};
};

namespace
{

int set_error(sw_ctx* ctx, int code, const std::string& msg)
{
	if(ctx)
		ctx->error = msg;
	return code;
}

int hip_check(sw_ctx* ctx, hipError_t e, const char* what)
{
	if(e == hipSuccess)
		return SW_OK;
	return set_error(ctx, SW_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int ensure(sw_ctx* ctx, DevBuf& b, size_t bytes)
{
	if(bytes <= b.cap && b.p)
		return SW_OK;
	size_t want = std::max<size_t>(bytes + bytes / 4, 256);
	if(b.p)
	{
		(void)hipStreamSynchronize(ctx->stream);
		(void)hipFree(b.p);
		ctx->device_bytes -= b.cap;
		b.p = nullptr;
		b.cap = 0;
	}
	void* p = nullptr;
	hipError_t e = hipMalloc(&p, want);
	if(e != hipSuccess)
	{
		(void)hipGetLastError();
		want = std::max<size_t>(bytes, 256);
		e = hipMalloc(&p, want);
		if(e != hipSuccess)
		{
			(void)hipGetLastError();
			return set_error(ctx, SW_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(want) + " bytes failed");
		}
	}
	b.p = p;
	b.cap = want;
	ctx->device_bytes += want;
	ctx->alloc_epoch++;
	return SW_OK;
}

void release(sw_ctx* ctx, DevBuf& b)
{
	if(b.p)
	{
		(void)hipFree(b.p);
		if(ctx)
			ctx->device_bytes -= b.cap;
	}
	b.p = nullptr;
	b.cap = 0;
}

int upload(sw_ctx* ctx, DevBuf& b, const void* src, size_t bytes)
{
	int r = ensure(ctx, b, bytes);
	if(r)
		return r;
	if(bytes == 0)
		return SW_OK;
	return hip_check(ctx, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync H2D");
}

uint64_t color_bytes(uint32_t bits) { return bits == 64 ? 8 : 4; }

enum
{
	STAGE_SETUP = 0,
	STAGE_SURF = 1,
	STAGE_RASTER = 2,
	STAGE_TONEMAP = 3,
	STAGE_SETUP_SURF = 4,
	STAGE_BLOOM = 5,
	STAGE_MODELS = 6,
	STAGE_TEXTURES = 7,
	STAGE_DEPTH = 8,
};

hipEvent_t take_event(sw_ctx* ctx)
{
	if(!ctx->event_pool.empty())
	{
		hipEvent_t e = ctx->event_pool.back();
		ctx->event_pool.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	if(hipEventCreate(&e) != hipSuccess)
	{
		(void)hipGetLastError();
		return nullptr;
	}
	return e;
}

// Brackets one launch with events on the context's stream when profiling is on.
// Times a stage with two events. The stage's first kernel launched through timed_launch records
// them at the kernel's own start and end; a stage without such a launch is timed by the events
// recorded on the stream around it (which adds a few microseconds of marker overhead).
struct StageTimer
{
	sw_ctx* ctx;
	int stage;
	hipEvent_t a = nullptr, b = nullptr;
	swk::LaunchEvents saved;
	StageTimer(sw_ctx* c, int s) : ctx(c), stage(s)
	{
		if(ctx->profiling && !ctx->capturing)
		{
			a = take_event(ctx);
			b = a ? take_event(ctx) : nullptr;
			if(a && !b)
			{
				ctx->event_pool.push_back(a);
				a = nullptr;
			}
			if(a)
			{
				// the start marker, overwritten by a timed launch's own start record
				(void)hipEventRecord(a, ctx->stream);
				saved = swk::launch_events();
				swk::launch_events() = swk::LaunchEvents{a, b, false};
			}
		}
	}
	~StageTimer()
	{
		if(!a)
			return;
		const bool attached = swk::launch_events().used;
		swk::launch_events() = saved;
		if(!attached)
			(void)hipEventRecord(b, ctx->stream); // no timed launch in the stage: the end marker
		ctx->pending.push_back(sw_ctx::Pending{stage, a, b});
	}
};

// Accumulates the elapsed times of completed launches (call after a stream synchronisation).
void collect_timings(sw_ctx* ctx)
{
	for(const sw_ctx::Pending& p : ctx->pending)
	{
		float ms = 0.0f;
		if(hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess)
		{
			ctx->prof_ms[p.stage] += ms;
			ctx->prof_count[p.stage]++;
		}
		else
			(void)hipGetLastError();
		ctx->event_pool.push_back(p.a);
		ctx->event_pool.push_back(p.b);
	}
	ctx->pending.clear();
}

// Per-view tiling of a batch (views are always host memory).
int build_view_ext(sw_ctx* ctx, const sw_view* views, uint32_t n, uint64_t arena_texels, std::vector<ViewExt>& out,
	uint32_t& total_tiles, uint32_t* max_height = nullptr)
{
	if(max_height)
		*max_height = 0;
	const uint32_t tile_rows = ctx->tile_rows;
	out.resize(n);
	uint64_t tiles = 0;
	uint32_t next_surface = 0, next_draw = 0;
	for(uint32_t i = 0; i < n; ++i)
	{
		const sw_view& v = views[i];
		if(v.width == 0 || v.height == 0 || v.width > 16384 || v.height > 16384)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "view " + std::to_string(i) + ": size must be in [1, 16384]");
		if(v.first_surface != next_surface || v.first_draw != next_draw)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT,
				"view " + std::to_string(i) + ": surfaces and draws of views must be consecutive ranges in view order");
		next_surface += v.num_surfaces;
		next_draw += v.num_draws;
		ViewExt& e = out[i];
		e.v = v;
		if(max_height)
			*max_height = std::max(*max_height, v.height);
		e.num_bands = (v.height + tile_rows - 1) / tile_rows;
		e.cols = (v.width + 63) / 64;
		e.first_tile = uint32_t(tiles);
		e.pad = 0;
		// a view's surfaces, and the arena texels its draws read, lie in [its offset, the next
		// view's offset) (the last view: to the arena's end)
		if(v.surface_arena_offset > arena_texels || (i > 0 && v.surface_arena_offset < views[i - 1].surface_arena_offset))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT,
				"view " + std::to_string(i) + ": surface arena offsets must be non-decreasing and inside the arena");
		e.arena_end = i + 1 < n ? std::max(views[i + 1].surface_arena_offset, v.surface_arena_offset) : arena_texels;
		tiles += uint64_t(e.num_bands) * e.cols;
		if(tiles >= (1ull << 31))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "batch too large");
	}
	total_tiles = uint32_t(tiles);
	return SW_OK;
}

int validate_views_fit(sw_ctx* ctx, const std::vector<ViewExt>& ve, uint64_t fb_texels, uint32_t num_surfaces, uint32_t num_draws)
{
	uint64_t ns = 0, nd = 0;
	for(size_t i = 0; i < ve.size(); ++i)
	{
		const sw_view& v = ve[i].v;
		if(v.fb_offset + uint64_t(v.width) * v.height > fb_texels)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "view " + std::to_string(i) + ": framebuffer range outside the output");
		ns += v.num_surfaces;
		nd += v.num_draws;
	}
	if(ns != num_surfaces || nd != num_draws)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "views do not cover the batch's surfaces and draws exactly");
	return SW_OK;
}

// Makes `ve` the views of the next launches: a single view travels in the kernel arguments
// (when allowed), several views are uploaded to the device array.
int set_views(sw_ctx* ctx, const std::vector<ViewExt>& ve, bool allow_inline)
{
	ViewsArg& va = ctx->views_arg;
	va.count = uint32_t(ve.size());
	if(allow_inline && ve.size() == 1)
	{
		va.inline_one = 1;
		va.one = ve[0];
		va.ptr = nullptr;
		return SW_OK;
	}
	int r = upload(ctx, ctx->views, ve.data(), sizeof(ViewExt) * ve.size());
	if(r)
		return r;
	va.inline_one = 0;
	va.one = ViewExt{};
	va.ptr = static_cast<const ViewExt*>(ctx->views.p);
	return SW_OK;
}

// Launch sequence of a batch: draw setup (with the surface extents when `surfaces`), surface
// building (when `surfaces`), tiled raster (when `raster`). The raster kernel's last block hands
// the counters over to the result block; a sequence without it must be followed by
// flush_surface_errors.
enum
{
	SURF_TABLE_DEVICE = 0,
	SURF_TABLE_HOST = 1,
	SURF_SEARCH = 2,
};

// Span records a batch may use with static per-draw storage (1 GiB of records); above it the
// draws allocate their records with atomics.
const uint64_t STATIC_SPAN_LIMIT = 1ull << 25;

int launch_stages(sw_ctx* ctx, const sw_ctx::Launch& L, bool surfaces, bool raster)
{
	sw_map* map = L.map;
	const uint32_t nd = raster ? L.num_draws : 0;
	const uint32_t ns = surfaces ? L.num_surfaces : 0;
	const uint64_t surf_bt = surf_block_texels(L.arena_texels);
	const uint64_t surf_blocks = (L.arena_texels + surf_bt - 1) / surf_bt;
	// a draw's span records never exceed its clip rect's height, at most the tallest view
	const uint64_t stride = L.max_view_height;
	const bool static_spans = nd > 0 && uint64_t(nd) * stride <= STATIC_SPAN_LIMIT;
	int r;
	if(static_spans && !ctx->capturing)
		ctx->span_capacity = std::max<uint64_t>(ctx->span_capacity, uint64_t(nd) * stride);
	if((r = ensure(ctx, ctx->parts, sizeof(PartRec) * uint64_t(std::max<uint32_t>(nd, 1)) * MAX_PARTS_PER_DRAW)))
		return r;
	if((r = ensure(ctx, ctx->spans, sizeof(SpanRec) * ctx->span_capacity)))
		return r;
	if((r = ensure(ctx, ctx->drec, sizeof(DrawRec) * std::max<uint32_t>(nd, 1))))
		return r;
	if((r = ensure(ctx, ctx->dbounds, sizeof(DrawBounds) * std::max<uint32_t>(nd, 1))))
		return r;
	if(L.surface_mode == SURF_TABLE_DEVICE)
	{
		if((r = ensure(ctx, ctx->surf_ext, sizeof(SurfExt) * std::max<uint32_t>(ns, 1))))
			return r;
		if((r = ensure(ctx, ctx->blk_surf, sizeof(uint32_t) * std::max<uint64_t>(surf_blocks, 1))))
			return r;
	}
	uint32_t* counters = static_cast<uint32_t*>(ctx->counters.p);
	MapPtrs mp = map->ptrs;
	mp.portal = L.portal;
	mp.num_portal = L.num_portal;
	SetupArgs a{};
	a.draws = L.draws;
	a.num_draws = nd;
	a.vdata = L.vdata;
	a.num_vertex_ints = L.num_vertex_ints;
	a.parts = static_cast<PartRec*>(ctx->parts.p);
	a.spans = static_cast<SpanRec*>(ctx->spans.p);
	a.drec = static_cast<DrawRec*>(ctx->drec.p);
	a.bounds = static_cast<DrawBounds*>(ctx->dbounds.p);
	a.host_err = ctx->counters_host_dev + HOST_ERR_OFFSET;
	a.counters = counters;
	a.span_capacity = ctx->span_capacity;
	a.span_stride = static_spans ? stride : 0;
	a.arena_texels = L.arena_texels;
	a.skybox_texels = L.color_bits == 64 ? map->num_sky64 : map->num_sky32;
	a.lite_texels = map->ptrs.num_lite;
	a.portal_texels = L.num_portal;
	SurfaceArgs sa{};
	sa.surfaces = L.surfaces;
	sa.num_surfaces = ns;
	sa.num_lights = L.num_lights;
	sa.surf_ext = static_cast<const SurfExt*>(ctx->surf_ext.p);
	sa.blk_surf = L.surface_mode == SURF_SEARCH ? nullptr : static_cast<const uint32_t*>(ctx->blk_surf.p);
	sa.counters = counters;
	sa.host_err = ctx->counters_host_dev + HOST_ERR_OFFSET;
	sa.lights = L.lights;
	sa.light_indices = L.light_indices;
	sa.num_light_indices = L.num_light_indices;
	sa.arena = L.arena;
	sa.arena_texels = L.arena_texels;
	sa.num_shadows = ctx->num_shadows;
	sa.shadows = static_cast<const sw_light_shadow*>(ctx->shadows.p);
	sa.shadow_texels = ctx->shadow_texels;
	// a block per draw while the draws alone cannot fill the chip, a wave per draw beyond
	const uint32_t lanes = nd <= 4096 ? 256u : 64u;
	if(surfaces && L.surface_mode != SURF_TABLE_DEVICE && ctx->num_shadows)
	{
		// shadowed lights: the surface kernel's shadow instantiation runs on its own after the
		// setup (whose table blocks are not needed: the host prepared the table, or the surface
		// blocks search)
		{
			StageTimer t(ctx, STAGE_SETUP);
			swk::launch_setup(ctx->stream, lanes, a, ctx->views_arg, L.surfaces, 0, static_cast<SurfExt*>(ctx->surf_ext.p),
				static_cast<uint32_t*>(ctx->blk_surf.p), mp);
		}
		StageTimer t(ctx, STAGE_SURF);
		swk::launch_build_surfaces(ctx->stream, L.color_bits, mp, sa, ctx->views_arg, counters);
	}
	else if(surfaces && L.surface_mode != SURF_TABLE_DEVICE)
	{
		StageTimer t(ctx, STAGE_SETUP_SURF);
		swk::launch_setup_and_surfaces(ctx->stream, lanes, L.color_bits, a, ctx->views_arg, sa, mp);
	}
	else
	{
		{
			StageTimer t(ctx, STAGE_SETUP);
			swk::launch_setup(ctx->stream, lanes, a, ctx->views_arg, L.surfaces, ns, static_cast<SurfExt*>(ctx->surf_ext.p),
				static_cast<uint32_t*>(ctx->blk_surf.p), mp);
		}
		if(surfaces)
		{
			StageTimer t(ctx, STAGE_SURF);
			swk::launch_build_surfaces(ctx->stream, L.color_bits, mp, sa, ctx->views_arg, counters);
		}
	}
	if(raster)
	{
		{
			StageTimer t(ctx, STAGE_RASTER);
			swk::launch_raster_tiles(ctx->stream, L.color_bits, ctx->tile_rows, ctx->scalar_spans, mp,
				static_cast<DrawBounds*>(ctx->dbounds.p),
				static_cast<DrawRec*>(ctx->drec.p), static_cast<PartRec*>(ctx->parts.p), static_cast<SpanRec*>(ctx->spans.p),
				ctx->views_arg, L.arena, L.color_bits == 64 ? map->sky64.p : map->sky32.p, L.fb, L.total_tiles, counters);
		}
		// With static span storage nothing in the live counters needs clearing after a sequence
		// (errors also reach the host words), so the statistics wait until they are asked for; the
		// span allocation counter needs k_finish to hand it over and clear it.
		ctx->finish_launched = !static_spans;
		ctx->stats_pending = static_spans;
		ctx->stats_draws = nd;
		if(!static_spans)
			swk::launch_finish(ctx->stream, static_cast<DrawRec*>(ctx->drec.p), nd, counters, ctx->counters_host_dev);
	}
	return hip_check(ctx, hipGetLastError(), "kernel launch");
}

// Host-side preparation of host-resident surface descriptors (what the setup launch's extra blocks
// do for device-resident ones): arena extents, validation and the block -> first surface table,
// uploaded for the surface kernel. Mirrors prep_surface in sw_kernels.hip.
int host_prepare_surfaces(sw_ctx* ctx, const sw_map* map, const std::vector<ViewExt>& ve, const sw_surface_desc* s, uint32_t ns,
	uint64_t arena_texels)
{
	std::vector<SurfExt>& ext = ctx->host_ext;
	std::vector<uint32_t>& blk = ctx->host_blk;
	const uint64_t bt = surf_block_texels(arena_texels);
	const uint64_t num_blocks = (arena_texels + bt - 1) / bt;
	ext.resize(std::max<uint32_t>(ns, 1));
	blk.resize(std::max<uint64_t>(num_blocks, 1));
	const std::vector<sw_texture_level>& lv = map->host_levels;
	uint32_t view = 0;
	uint64_t prev_end = 0;
	for(uint32_t i = 0; i < ns; ++i)
	{
		while(view + 1 < ve.size() && ve[view + 1].v.first_surface <= i)
			++view;
		const sw_surface_desc& d = s[i];
		const uint64_t start = (ve.empty() ? 0 : ve[view].v.surface_arena_offset) + d.pixels_offset;
		const uint64_t end = start + uint64_t(d.size[0]) * d.size[1];
		const uint64_t limit = ve.empty() ? arena_texels : ve[view].arena_end;
		bool bad = end > limit || d.size[0] > 65536u || d.size[1] > 65536u || start < prev_end;
		if((d.flags & SW_SURF_DIFFUSE_LAYER) && (d.texture_level >= lv.size() || (lv[d.texture_level].flags & SW_TEX_LITE_LEVEL)))
			bad = true;
		if((d.flags & SW_SURF_EMISSIVE) && d.emissive_level != SW_NO_TEXTURE &&
			(d.emissive_level >= lv.size() || !(lv[d.emissive_level].flags & SW_TEX_LITE_LEVEL)))
			bad = true;
		if((d.flags & SW_SURF_DIFFUSE_LAYER) && d.lightmap_offset != 0 && d.size[0] > 0 && d.size[1] > 0)
		{
			const uint32_t sl = d.lightmap_scale_log2;
			const uint64_t num_lm = (d.flags & SW_SURF_DIRECTIONAL) ? map->ptrs.num_lm_dir : map->ptrs.num_lm_simple;
			if(sl > 4 || d.lightmap_size[0] < 2 || d.lightmap_size[1] < 2 ||
				((uint64_t(d.size[0]) - 1 + d.lightmap_tc_shift[0]) >> sl) + 1 >= d.lightmap_size[0] ||
				((uint64_t(d.size[1]) - 1 + d.lightmap_tc_shift[1]) >> sl) + 1 >= d.lightmap_size[1] ||
				uint64_t(d.lightmap_offset) + uint64_t(d.lightmap_size[0]) * d.lightmap_size[1] > num_lm)
				bad = true;
		}
		if(bad)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT,
				"surface " + std::to_string(i) +
					": invalid descriptor (arena order/overlap, texture level, or lightmap range out of bounds)");
		ext[i] = SurfExt{end, view, 0u};
		const uint64_t b0 = (prev_end + bt - 1) / bt;
		const uint64_t b1 = i + 1 == ns ? num_blocks : std::min<uint64_t>((end + bt - 1) / bt, num_blocks);
		for(uint64_t b = b0; b < b1; ++b)
			blk[b] = i;
		prev_end = end;
	}
	int r = upload(ctx, ctx->surf_ext, ext.data(), sizeof(SurfExt) * ext.size());
	return r ? r : upload(ctx, ctx->blk_surf, blk.data(), sizeof(uint32_t) * blk.size());
}

// The result block k_finish wrote into host memory -> the pixel statistics.
void read_result_block(sw_ctx* ctx)
{
	uint32_t c[COUNTER_COUNT];
	for(int k = 0; k < COUNTER_COUNT; ++k)
		c[k] = static_cast<volatile const uint32_t*>(ctx->counters_host)[k];
	auto u64 = [&](int lo) { return uint64_t(c[lo]) | (uint64_t(c[lo + 1]) << 32); };
	ctx->stats.polygon_pixels = u64(COUNTER_POLY_PIXELS_LO);
	ctx->stats.triangle_pixels = u64(COUNTER_TRI_PIXELS_LO);
	ctx->stats.blend_pixels = u64(COUNTER_BLEND_PIXELS_LO);
	ctx->stats.spans = u64(COUNTER_SPANS_LO);
	ctx->stats.parts = c[COUNTER_PARTS];
}

// Computes the last raster batch's statistics if they are still pending (one k_finish over its
// draw records, which also clears the live counters).
int ensure_stats(sw_ctx* ctx)
{
	if(!ctx->stats_pending)
		return SW_OK;
	ctx->stats_pending = false;
	swk::launch_finish(ctx->stream, static_cast<DrawRec*>(ctx->drec.p), ctx->stats_draws,
		static_cast<uint32_t*>(ctx->counters.p), ctx->counters_host_dev);
	int r = hip_check(ctx, hipGetLastError(), "kernel launch");
	if(!r)
		r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
	if(r)
		return r;
	read_result_block(ctx);
	return SW_OK;
}

// Reads and clears the error words the kernels of the last sequence set in host memory; after an
// error the live counters are cleared as well (a sequence without errors leaves them clear).
uint32_t take_errors(sw_ctx* ctx)
{
	volatile uint32_t* w = ctx->counters_host + HOST_ERR_OFFSET;
	uint32_t flags = 0;
	for(int k = 0; k < HOST_ERR_WORDS; ++k)
		if(w[k])
		{
			flags |= 1u << k;
			w[k] = 0;
		}
	if(flags)
		hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, ctx->stream, static_cast<uint32_t*>(ctx->counters.p));
	return flags;
}

// After a surface-only launch sequence: waits, keeps its error flag, clears the live counters.
int flush_surface_errors(sw_ctx* ctx)
{
	uint32_t* counters = static_cast<uint32_t*>(ctx->counters.p);
	int r = hip_check(ctx,
		hipMemcpyAsync(ctx->counters_host, counters, sizeof(uint32_t) * COUNTER_COUNT, hipMemcpyDeviceToHost, ctx->stream),
		"hipMemcpyAsync counters");
	if(r)
		return r;
	hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, ctx->stream, counters);
	if((r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")))
		return r;
	collect_timings(ctx);
	if((ctx->counters_host[COUNTER_ERROR] & ERROR_BAD_SURFACE) | (take_errors(ctx) & ERROR_BAD_SURFACE))
		ctx->surface_error = true;
	return SW_OK;
}

// Waits for the batch, reads the counters; re-runs the raster stages if span storage overflowed.
int finish(sw_ctx* ctx)
{
	ctx->stats.reruns = 0;
	for(int attempt = 0; attempt < 4; ++attempt)
	{
		int r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
		if(r)
			return r;
		collect_timings(ctx);
		const uint32_t errors = take_errors(ctx);
		const bool surface_error = ctx->surface_error;
		ctx->surface_error = false;
		if(surface_error || (errors & ERROR_BAD_SURFACE))
		{
			ctx->stats_pending = false;
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT,
				"invalid surface descriptor (arena order/overlap, texture level, or lightmap range out of bounds)");
		}
		if(errors & ERROR_PART_GAP)
		{
			ctx->stats_pending = false;
			return set_error(ctx, SW_ERR_DEVICE, "internal: a draw's parts did not cover consecutive rows");
		}
		if(ctx->finish_launched)
			read_result_block(ctx); // k_finish ran: the statistics (and the span count) are in
		if(errors & ERROR_SPAN_OVERFLOW)
		{
			// only span allocation overflows, and it always ends with k_finish
			const uint64_t needed = ctx->stats.spans;
			if(needed >= (1ull << 31))
				return set_error(ctx, SW_ERR_BUDGET, "span storage above 2^31 records");
			ctx->span_capacity = std::max<uint64_t>(needed + needed / 4, ctx->span_capacity * 2);
			if((r = launch_stages(ctx, ctx->last, false, true)))
				return r;
			ctx->stats.reruns++;
			continue;
		}
		return SW_OK;
	}
	return set_error(ctx, SW_ERR_BUDGET, "span storage overflow persisted");
}

void fill_surface_stats(sw_ctx* ctx, const sw_surface_desc* s, uint32_t n, bool host)
{
	ctx->stats.surface_texels = 0;
	ctx->stats.lightmap_elements = 0;
	ctx->stats.emissive_texels = 0;
	if(!host)
		return;
	for(uint32_t i = 0; i < n; ++i)
	{
		const uint64_t t = uint64_t(s[i].size[0]) * s[i].size[1];
		if(s[i].flags & SW_SURF_DIFFUSE_LAYER)
		{
			ctx->stats.surface_texels += t;
			if(s[i].lightmap_offset)
				ctx->stats.lightmap_elements += 2ull * s[i].lightmap_size[0] * s[i].size[1];
		}
		if(s[i].flags & SW_SURF_EMISSIVE)
			ctx->stats.emissive_texels += t;
	}
}

// Launch sequence of the plain path's frame: the pending surfaces (if any), then the draws
// already uploaded to ctx->draws / ctx->vdata when `raster`.
int launch_frame(sw_ctx* ctx, uint32_t num_draws, uint64_t num_vertex_ints, bool raster)
{
	int r;
	if(ctx->arena.p == nullptr && (r = ensure(ctx, ctx->arena, 8)))
		return r;
	sw_view v = ctx->frame_view;
	v.first_surface = 0;
	v.num_surfaces = ctx->frame_num_surfaces;
	v.first_draw = 0;
	v.num_draws = raster ? num_draws : 0;
	v.fb_offset = 0;
	v.surface_arena_offset = 0;
	std::vector<ViewExt> ve;
	sw_ctx::Launch& L = ctx->last;
	L = sw_ctx::Launch{};
	if((r = build_view_ext(ctx, &v, 1, ctx->frame_arena_texels, ve, L.total_tiles, &L.max_view_height)))
		return r;
	L.surface_mode = SURF_SEARCH;
	if((r = set_views(ctx, ve, true)))
		return r;
	L.map = ctx->frame_map;
	L.color_bits = ctx->frame_color_bits;
	L.num_views = 1;
	L.surfaces = static_cast<const sw_surface_desc*>(ctx->surfaces.p);
	L.num_surfaces = ctx->frame_num_surfaces;
	L.lights = static_cast<const sw_light*>(ctx->lights.p);
	L.num_lights = ctx->frame_num_lights;
	L.light_indices = static_cast<const uint32_t*>(ctx->light_indices.p);
	L.num_light_indices = ctx->frame_num_light_indices;
	L.draws = static_cast<const sw_draw*>(ctx->draws.p);
	L.num_draws = raster ? num_draws : 0;
	L.vdata = static_cast<const int32_t*>(ctx->vdata.p);
	L.num_vertex_ints = raster ? num_vertex_ints : 0;
	L.arena_texels = ctx->frame_arena_texels;
	L.arena = ctx->arena.p;
	L.fb = ctx->fb.p;
	L.portal = ctx->portal_texels;
	L.num_portal = ctx->portal_texels ? ctx->num_portal_texels : 0;
	const bool surfaces = ctx->frame_surfaces_pending;
	ctx->frame_surfaces_pending = false;
	if(!raster && !surfaces)
		return SW_OK;
	if((r = launch_stages(ctx, L, surfaces, raster)))
		return r;
	return raster ? SW_OK : flush_surface_errors(ctx);
}

} // namespace

// =============================================================================================
// C ABI
// =============================================================================================
namespace
{
const int MAX_SCRATCH_DEVICES = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[MAX_SCRATCH_DEVICES] = {};

// The device address of `scratch` on `device`, which must be current: looked up once per device
// (each GPU holds its own copy of the array, at its own address).
int scratch_address(int device, float** out)
{
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if(device < 0 || device >= MAX_SCRATCH_DEVICES)
		return SW_ERR_DEVICE;
	if(!g_scratch_ptr[device])
	{
		void* p = nullptr;
		if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
		{
			(void)hipGetLastError();
			return SW_ERR_DEVICE;
		}
		g_scratch_ptr[device] = static_cast<float*>(p);
	}
	*out = g_scratch_ptr[device];
	return SW_OK;
}
} // namespace

extern "C" {

int sw_ctx_create(int device, sw_ctx** out_ctx)
{
	if(!out_ctx)
		return SW_ERR_INVALID_ARGUMENT;
	*out_ctx = nullptr;
	int count = 0;
	if(hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
	{
		(void)hipGetLastError();
		return SW_ERR_DEVICE;
	}
	sw_ctx* ctx = new(std::nothrow) sw_ctx();
	if(!ctx)
		return SW_ERR_OUT_OF_MEMORY;
	ctx->device = device;
	if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
	{
		(void)hipGetLastError();
		delete ctx;
		return SW_ERR_DEVICE;
	}
	if(ensure(ctx, ctx->counters, sizeof(uint32_t) * COUNTER_TOTAL) ||
		hipMemsetAsync(ctx->counters.p, 0, sizeof(uint32_t) * COUNTER_TOTAL, ctx->stream) != hipSuccess ||
		hipStreamSynchronize(ctx->stream) != hipSuccess ||
		hipHostMalloc(reinterpret_cast<void**>(&ctx->counters_host), sizeof(uint32_t) * HOST_WORDS_TOTAL,
			hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
		hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->counters_host_dev), ctx->counters_host, 0) != hipSuccess)
	{
		sw_ctx_destroy(ctx);
		return SW_ERR_DEVICE;
	}
	std::memset(ctx->counters_host, 0, sizeof(uint32_t) * HOST_WORDS_TOTAL);
	if(scratch_address(device, &ctx->noalloc_scratch))
	{
		sw_ctx_destroy(ctx);
		return SW_ERR_DEVICE;
	}
	*out_ctx = ctx;
	return SW_OK;
}

void sw_ctx_destroy(sw_ctx* ctx)
{
	if(!ctx)
		return;
	(void)hipSetDevice(ctx->device);
	if(ctx->stream)
		(void)hipStreamSynchronize(ctx->stream);
	if(ctx->graph_exec)
		(void)hipGraphExecDestroy(ctx->graph_exec);
	if(ctx->graph)
		(void)hipGraphDestroy(ctx->graph);
	DevBuf* bufs[] = {&ctx->views, &ctx->surfaces, &ctx->lights, &ctx->light_indices, &ctx->draws, &ctx->vdata, &ctx->surf_ext, &ctx->blk_surf,
		&ctx->parts, &ctx->spans, &ctx->drec, &ctx->dbounds, &ctx->counters, &ctx->arena, &ctx->fb, &ctx->tonemap_out, &ctx->tonemap_sums,
		&ctx->bloom0, &ctx->bloom1, &ctx->bloom_lines, &ctx->model_args, &ctx->model_scratch, &ctx->texop_args, &ctx->shadows, &ctx->depth_args, &ctx->depth_vdata,
		&ctx->depth_rows, &ctx->depth_draw_rows, &ctx->water_points, &ctx->fire_args};
	for(DevBuf* b : bufs)
		release(ctx, *b);
	if(ctx->counters_host)
		(void)hipHostFree(ctx->counters_host);
	for(uint32_t k = 0; k < ARG_STAGING_SLOTS; ++k)
	{
		if(ctx->arg_staging[k])
			(void)hipHostFree(ctx->arg_staging[k]);
		if(ctx->arg_staging_event[k])
			(void)hipEventDestroy(ctx->arg_staging_event[k]);
	}
	collect_timings(ctx);
	for(hipEvent_t e : ctx->event_pool)
		(void)hipEventDestroy(e);
	if(ctx->stream)
		(void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

const char* sw_last_error(const sw_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int sw_sync(sw_ctx* ctx)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	const int r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
	if(r)
		return r;
	volatile uint32_t* model_err = ctx->counters_host + HOST_MODEL_ERR_OFFSET;
	if(*model_err)
	{
		*model_err = 0;
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT,
			"model preparation: a triangle's vertex index, a vertex's bone index or a bone's parent index was out of range");
	}
	return SW_OK;
}

void* sw_ctx_stream(sw_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

uint64_t sw_ctx_device_bytes(const sw_ctx* ctx) { return ctx ? ctx->device_bytes : 0; }

int sw_map_upload(sw_ctx* ctx, const sw_map_desc* d, sw_map** out_map)
{
	if(!ctx || !d || !out_map)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	*out_map = nullptr;
	(void)hipSetDevice(ctx->device);
	// The kernels address each buffer with 32-bit byte offsets from its base.
	const uint64_t lim = 1ull << 32;
	if(d->num_texels >= lim / sizeof(sw_texel) || d->num_lite_texels >= lim / sizeof(uint32_t) ||
		d->num_lightmap_elements >= lim / (3 * sizeof(float)) ||
		d->num_directional_lightmap_elements >= lim / sizeof(sw_directional_lightmap_element))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "map buffers above 4 GiB each (texels, lite texels, lightmaps)");
	// Validate the texture table against its buffers.
	for(uint32_t i = 0; i < d->num_levels; ++i)
	{
		const sw_texture_level& l = d->levels[i];
		const uint64_t n = (l.flags & SW_TEX_LITE_LEVEL) ? d->num_lite_texels : d->num_texels;
		if(l.width == 0 || l.height == 0 || l.width > 65536 || l.height > 65536 || l.offset + uint64_t(l.width) * l.height > n)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "texture level " + std::to_string(i) + " outside its texel buffer");
	}
	sw_map* m = new(std::nothrow) sw_map();
	if(!m)
		return set_error(ctx, SW_ERR_OUT_OF_MEMORY, "map allocation");
	int r = SW_OK;
	r = r ? r : upload(ctx, m->texels, d->texels, sizeof(sw_texel) * d->num_texels);
	r = r ? r : upload(ctx, m->lite, d->lite_texels, sizeof(uint32_t) * d->num_lite_texels);
	// Both lightmap kinds in one buffer (directional elements first, the simple ones at a 256-byte
	// aligned offset), so that the kernels address either with one base and a 32-bit offset.
	const uint64_t dir_bytes = sizeof(sw_directional_lightmap_element) * d->num_directional_lightmap_elements;
	const uint64_t simple_off = (dir_bytes + 255) & ~uint64_t(255);
	const uint64_t simple_bytes = sizeof(float) * 3 * d->num_lightmap_elements;
	if(simple_off + simple_bytes >= (1ull << 32))
		r = set_error(ctx, SW_ERR_INVALID_ARGUMENT, "lightmaps above 4 GiB");
	r = r ? r : ensure(ctx, m->lm_dir, simple_off + simple_bytes);
	if(!r && dir_bytes)
		r = hip_check(ctx, hipMemcpyAsync(m->lm_dir.p, d->lightmaps_directional, dir_bytes, hipMemcpyHostToDevice, ctx->stream),
			"hipMemcpyAsync H2D");
	if(!r && simple_bytes)
		r = hip_check(ctx,
			hipMemcpyAsync(static_cast<char*>(m->lm_dir.p) + simple_off, d->lightmaps_simple, simple_bytes, hipMemcpyHostToDevice,
				ctx->stream),
			"hipMemcpyAsync H2D");
	r = r ? r : upload(ctx, m->levels, d->levels, sizeof(sw_texture_level) * d->num_levels);
	r = r ? r : upload(ctx, m->sky32, d->skybox_texels32, sizeof(uint32_t) * (d->skybox_texels32 ? d->num_skybox_texels32 : 0));
	r = r ? r : upload(ctx, m->sky64, d->skybox_texels64, sizeof(uint64_t) * (d->skybox_texels64 ? d->num_skybox_texels64 : 0));
	if(!r)
		r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "map upload");
	if(r)
	{
		sw_map_release(ctx, m);
		return r;
	}
	m->num_sky32 = d->skybox_texels32 ? d->num_skybox_texels32 : 0;
	m->num_sky64 = d->skybox_texels64 ? d->num_skybox_texels64 : 0;
	m->host_levels.assign(d->levels, d->levels + d->num_levels);
	m->ptrs.texels = static_cast<const sw_texel*>(m->texels.p);
	m->ptrs.lite = static_cast<const uint32_t*>(m->lite.p);
	m->ptrs.lm_dir = static_cast<const sw_directional_lightmap_element*>(m->lm_dir.p);
	m->ptrs.lm_simple = reinterpret_cast<const float*>(static_cast<const char*>(m->lm_dir.p) + simple_off);
	m->ptrs.lm_simple_offset = uint32_t(simple_off);
	m->ptrs.levels = static_cast<const sw_texture_level*>(m->levels.p);
	m->ptrs.num_levels = d->num_levels;
	m->ptrs.num_texels = d->num_texels;
	m->ptrs.num_lite = d->num_lite_texels;
	m->ptrs.num_lm_simple = d->num_lightmap_elements;
	m->ptrs.num_lm_dir = d->num_directional_lightmap_elements;
	*out_map = m;
	return SW_OK;
}

void sw_map_release(sw_ctx* ctx, sw_map* m)
{
	if(!m)
		return;
	if(ctx)
		(void)hipStreamSynchronize(ctx->stream);
	DevBuf* bufs[] = {&m->texels, &m->lite, &m->lm_simple, &m->lm_dir, &m->levels, &m->sky32, &m->sky64};
	for(DevBuf* b : bufs)
		release(ctx, *b);
	delete m;
}

// ---------------------------------------------------------------------------------------------
// Plain path
// ---------------------------------------------------------------------------------------------
int sw_frame_begin(sw_ctx* ctx, sw_map* map, uint32_t color_bits, uint32_t width, uint32_t height, const float cam_pos[3],
	uint32_t clear_background, uint64_t background_color)
{
	if(!ctx || !map || !cam_pos)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(color_bits != 32 && color_bits != 64)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "color_bits must be 32 or 64");
	if(width == 0 || height == 0 || width > 16384 || height > 16384)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "frame size must be in [1, 16384]");
	(void)hipSetDevice(ctx->device);
	int r = ensure(ctx, ctx->fb, uint64_t(width) * height * color_bytes(color_bits));
	if(r)
		return r;
	ctx->graph_key = 0;
	ctx->frame_map = map;
	ctx->frame_color_bits = color_bits;
	sw_view& v = ctx->frame_view;
	v = sw_view{};
	v.width = width;
	v.height = height;
	v.cam_pos[0] = cam_pos[0];
	v.cam_pos[1] = cam_pos[1];
	v.cam_pos[2] = cam_pos[2];
	v.clear_background = clear_background;
	v.background_color = background_color;
	ctx->frame_arena_texels = 0;
	ctx->frame_num_surfaces = 0;
	ctx->frame_num_lights = 0;
	ctx->frame_num_light_indices = 0;
	ctx->frame_surfaces_pending = false;
	ctx->surface_error = false;
	ctx->frame_active = true;
	ctx->frame_rastered = false;
	return SW_OK;
}

int sw_build_surfaces(sw_ctx* ctx, const sw_surface_desc* surfaces, uint32_t num_surfaces, const sw_light* lights,
	uint32_t num_lights, const uint32_t* light_indices, uint32_t num_light_indices)
{
	if(!ctx || !ctx->frame_active)
		return set_error(ctx, SW_ERR_STATE, "sw_build_surfaces outside sw_frame_begin");
	if(num_surfaces && !surfaces)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null surfaces");
	(void)hipSetDevice(ctx->device);
	uint64_t arena = 0;
	for(uint32_t i = 0; i < num_surfaces; ++i)
		arena = std::max<uint64_t>(arena, surfaces[i].pixels_offset + uint64_t(surfaces[i].size[0]) * surfaces[i].size[1]);
	int r;
	if((r = ensure(ctx, ctx->arena, std::max<uint64_t>(arena, 1) * color_bytes(ctx->frame_color_bits))))
		return r;
	if((r = upload(ctx, ctx->surfaces, surfaces, sizeof(sw_surface_desc) * num_surfaces)))
		return r;
	if((r = upload(ctx, ctx->lights, lights, sizeof(sw_light) * (lights ? num_lights : 0))))
		return r;
	if((r = upload(ctx, ctx->light_indices, light_indices, sizeof(uint32_t) * (light_indices ? num_light_indices : 0))))
		return r;
	ctx->frame_arena_texels = arena;
	ctx->last_color_bits = ctx->frame_color_bits;
	ctx->frame_num_surfaces = num_surfaces;
	ctx->frame_num_lights = lights ? num_lights : 0;
	ctx->frame_num_light_indices = light_indices ? num_light_indices : 0;
	fill_surface_stats(ctx, surfaces, num_surfaces, true);
	// built by the launch sequence of sw_raster_draws / sw_frame_end (or sw_read_surfaces)
	ctx->frame_surfaces_pending = true;
	return SW_OK;
}

int sw_set_portal_textures(sw_ctx* ctx, const void* portal_texels_device, uint64_t num_portal_texels)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!portal_texels_device && num_portal_texels)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null portal textures");
	ctx->portal_texels = portal_texels_device;
	ctx->num_portal_texels = portal_texels_device ? num_portal_texels : 0;
	ctx->graph_key = 0;
	return SW_OK;
}

int sw_raster_draws(sw_ctx* ctx, const sw_draw* draws, uint32_t num_draws, const int32_t* vertex_data, uint64_t num_vertex_ints)
{
	if(!ctx || !ctx->frame_active)
		return set_error(ctx, SW_ERR_STATE, "sw_raster_draws outside sw_frame_begin");
	if((num_draws && !draws) || (num_vertex_ints && !vertex_data))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null draws / vertex data");
	(void)hipSetDevice(ctx->device);
	int r;
	if((r = upload(ctx, ctx->draws, draws, sizeof(sw_draw) * num_draws)))
		return r;
	if((r = upload(ctx, ctx->vdata, vertex_data, sizeof(int32_t) * num_vertex_ints)))
		return r;
	if((r = launch_frame(ctx, num_draws, num_vertex_ints, true)))
		return r;
	ctx->frame_rastered = true;
	return SW_OK;
}

int sw_frame_end(sw_ctx* ctx, void* dst, size_t dst_pitch_bytes, uint32_t dst_on_device)
{
	if(!ctx || !ctx->frame_active)
		return set_error(ctx, SW_ERR_STATE, "sw_frame_end outside sw_frame_begin");
	(void)hipSetDevice(ctx->device);
	int r;
	if(!ctx->frame_rastered)
	{
		// no draws: the framebuffer is the cleared background (or untouched)
		if((r = sw_raster_draws(ctx, nullptr, 0, nullptr, 0)))
			return r;
	}
	if((r = finish(ctx)))
		return r;
	if(dst)
	{
		const size_t row = size_t(ctx->frame_view.width) * color_bytes(ctx->frame_color_bits);
		if(dst_pitch_bytes < row)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "dst pitch smaller than a row");
		if((r = hip_check(ctx,
				hipMemcpy2DAsync(dst, dst_pitch_bytes, ctx->fb.p, row, row, ctx->frame_view.height,
					dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream),
				"hipMemcpy2DAsync framebuffer")))
			return r;
		if((r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")))
			return r;
	}
	ctx->frame_active = false;
	return SW_OK;
}

int sw_tonemap(sw_ctx* ctx, float exposure, const float color_modulate[3], void* dst32, size_t dst_pitch_bytes,
	uint32_t dst_on_device, float out_colors_sum[3])
{
	if(!ctx || !color_modulate)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(ctx->frame_color_bits != 64 || !ctx->fb.p)
		return set_error(ctx, SW_ERR_STATE, "sw_tonemap needs a 64-bit (HDR) frame");
	(void)hipSetDevice(ctx->device);
	int r;
	if(ctx->frame_active)
	{
		if(!ctx->frame_rastered && (r = sw_raster_draws(ctx, nullptr, 0, nullptr, 0)))
			return r;
		if((r = finish(ctx)))
			return r;
		ctx->frame_active = false;
	}
	const uint32_t w = ctx->frame_view.width, h = ctx->frame_view.height;
	void* out = dst32;
	uint32_t pitch_px = uint32_t(dst_pitch_bytes / 4);
	if(!dst_on_device)
	{
		if((r = ensure(ctx, ctx->tonemap_out, uint64_t(w) * h * 4)))
			return r;
		out = ctx->tonemap_out.p;
		pitch_px = w;
	}
	else if(dst_pitch_bytes < size_t(w) * 4 || dst_pitch_bytes % 4)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "bad dst pitch");
	if((r = sw_tonemap_device(ctx, ctx->fb.p, w, h, exposure, color_modulate, out, size_t(pitch_px) * 4, out_colors_sum)))
		return r;
	if(!dst_on_device && dst32)
	{
		if(dst_pitch_bytes < size_t(w) * 4)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "bad dst pitch");
		if((r = hip_check(ctx,
				hipMemcpy2DAsync(dst32, dst_pitch_bytes, out, size_t(w) * 4, size_t(w) * 4, h, hipMemcpyDeviceToHost, ctx->stream),
				"hipMemcpy2DAsync tonemap")))
			return r;
		return hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
	}
	return SW_OK;
}

int sw_tonemap_device(sw_ctx* ctx, const void* hdr_device, uint32_t width, uint32_t height, float exposure,
	const float color_modulate[3], void* dst32_device, size_t dst_pitch_bytes, float out_colors_sum[3])
{
	if(!ctx || !hdr_device || !dst32_device || !color_modulate)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(dst_pitch_bytes < size_t(width) * 4 || dst_pitch_bytes % 4)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "bad dst pitch");
	(void)hipSetDevice(ctx->device);
	int r;
	if((r = ensure(ctx, ctx->tonemap_sums, sizeof(float) * 3 * swk::TONEMAP_BLOCKS)))
		return r;
	float* host_sums = reinterpret_cast<float*>(ctx->counters_host + HOST_POST_OFFSET);
	float* host_sums_dev = reinterpret_cast<float*>(ctx->counters_host_dev + HOST_POST_OFFSET);
	// TonemappingFunction::new (postprocessor.rs:1005-1017)
	const float isr = 1.0f / (exposure * color_modulate[0]);
	const float isg = 1.0f / (exposure * color_modulate[1]);
	const float isb = 1.0f / (exposure * color_modulate[2]);
	{
		StageTimer t(ctx, STAGE_TONEMAP);
		swk::launch_tonemap(ctx->stream, static_cast<const uint64_t*>(hdr_device), width, height, isr, isg, isb,
			static_cast<uint32_t*>(dst32_device), uint32_t(dst_pitch_bytes / 4), static_cast<float*>(ctx->tonemap_sums.p),
			out_colors_sum ? host_sums_dev : nullptr);
	}
	if((r = hip_check(ctx, hipGetLastError(), "tonemap launch")))
		return r;
	if(out_colors_sum)
	{
		// the sums kernel stores the totals straight into the pinned host words
		if((r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")))
			return r;
		for(int k = 0; k < 3; ++k)
			out_colors_sum[k] = static_cast<volatile const float*>(host_sums)[k];
	}
	return SW_OK;
}

int sw_tonemap_device_noalloc(sw_ctx* ctx, const void* hdr_device, uint32_t width, uint32_t height, float exposure,
	const float color_modulate[3], void* dst32_device, size_t dst_pitch_bytes, float* out_colors_sum_device)
{
	if(!ctx || !hdr_device || !dst32_device || !color_modulate)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(dst_pitch_bytes < size_t(width) * 4 || dst_pitch_bytes % 4 || dst_pitch_bytes / 4 > 0xFFFFFFFFull)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "bad dst pitch");
	// the launch's blocks each leave three partial sums in the static scratch
	const uint64_t partial_floats = 3ull * std::max(1u, std::min(height, swk::TONEMAP_BLOCKS));
	if(partial_floats > sizeof(scratch) / sizeof(float))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "tonemap partial sums exceed the static scratch (3 * 2048 floats)");
	if(!ctx->noalloc_scratch)
		return set_error(ctx, SW_ERR_STATE, "no scratch address for this device");
	(void)hipSetDevice(ctx->device);
	// TonemappingFunction::new (postprocessor.rs:1005-1017)
	const float isr = 1.0f / (exposure * color_modulate[0]);
	const float isg = 1.0f / (exposure * color_modulate[1]);
	const float isb = 1.0f / (exposure * color_modulate[2]);
	// every partial the launch uses is written before k_tonemap_sums reads it: nothing to clear
	swk::launch_tonemap(ctx->stream, static_cast<const uint64_t*>(hdr_device), width, height, isr, isg, isb,
		static_cast<uint32_t*>(dst32_device), uint32_t(dst_pitch_bytes / 4), ctx->noalloc_scratch, out_colors_sum_device);
	return hip_check(ctx, hipGetLastError(), "tonemap launch");
}

// compute_gaussian_kernel (postprocessor.rs:1026-1045): four samples per tap, f32, libm exp
static void bloom_kernel(float sigma, int radius, float out[BLOOM_KERNEL])
{
	for(int i = 0; i < BLOOM_KERNEL; ++i)
		out[i] = 0.0f;
	const float samples[4] = {-0.375f, -0.125f, 0.125f, 0.375f};
	const float sqrt_tau = std::sqrt(6.28318530717958647692f);
	for(int x = -radius; x <= radius; ++x)
	{
		float val = 0.0f;
		for(float sample : samples)
		{
			const float coord = float(x) + sample;
			val += std::exp(-0.5f * (coord / sigma) * (coord / sigma)) / (sigma * sqrt_tau);
		}
		out[x + radius] = val / 4.0f;
	}
}

int sw_postprocess_device(sw_ctx* ctx, const void* hdr_device, uint32_t width, uint32_t height, float exposure,
	const float color_modulate[3], const sw_bloom_params* bloom, void* dst32_device, size_t dst_pitch_bytes,
	float out_average_rgb[3])
{
	if(!ctx || !hdr_device || !dst32_device || !color_modulate)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(dst_pitch_bytes < size_t(width) * 4 || dst_pitch_bytes % 4)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "bad dst pitch");
	if(height > 65535)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "frame taller than 65535 rows");
	(void)hipSetDevice(ctx->device);
	// postprocessor.rs:928-940 clamps the settings
	const float sigma = bloom ? std::min(std::max(bloom->sigma, 0.0f), 40.0f) : 0.0f;
	if(!(sigma > 0.0f))
	{
		float sums[3] = {0.0f, 0.0f, 0.0f};
		const int r = sw_tonemap_device(ctx, hdr_device, width, height, exposure, color_modulate, dst32_device, dst_pitch_bytes,
			out_average_rgb ? sums : nullptr);
		if(r == SW_OK && out_average_rgb)
		{
			// postprocessor.rs:181-184
			const float inv = 1.0f / float(uint64_t(width) * height);
			for(int k = 0; k < 3; ++k)
				out_average_rgb[k] = sums[k] * inv;
		}
		return r;
	}
	BloomArgs a{};
	a.scale_log2 = std::min<uint32_t>(std::max<uint32_t>(bloom->scale_log2, 1u), 3u);
	const float bloom_scale = std::min(std::max(bloom->scale, 1.0f / 16.0f), 0.5f);
	const uint32_t S = 1u << a.scale_log2;
	a.hdr = static_cast<const uint64_t*>(hdr_device);
	a.w = width;
	a.h = height;
	a.bw = width / S;
	a.bh = height / S;
	if(a.bw == 0 || a.bh == 0)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "frame smaller than one bloom texel");
	// perform_bloom (postprocessor.rs:728-803): radius and integer weights of both passes
	const float sigma_c = sigma / float(S);
	a.radius = std::min(int32_t(std::max(std::ceil(3.0f * sigma_c - 0.5f), 0.0f)), int32_t(BLOOM_MAX_RADIUS));
	float kernel[BLOOM_KERNEL];
	bloom_kernel(sigma_c, a.radius, kernel);
	for(int i = 0; i < BLOOM_KERNEL; ++i)
	{
		a.k1[i] = int32_t(kernel[i] * float(1 << 14));
		a.k2[i] = int32_t(kernel[i] * bloom_scale * float(1 << 14));
	}
	// TonemappingFunction::new (postprocessor.rs:1005-1017), lanes B, G, R
	a.inv_scale[0] = 1.0f / (exposure * color_modulate[2]);
	a.inv_scale[1] = 1.0f / (exposure * color_modulate[1]);
	a.inv_scale[2] = 1.0f / (exposure * color_modulate[0]);
	a.linear = bloom->linear_filter ? 1u : 0u;
	a.dst = static_cast<uint32_t*>(dst32_device);
	a.dst_pitch = uint32_t(dst_pitch_bytes / 4);
	const uint64_t texels = uint64_t(a.bw) * a.bh;
	int r;
	if((r = ensure(ctx, ctx->bloom0, texels * 8)) || (r = ensure(ctx, ctx->bloom1, texels * 8)) ||
		(r = ensure(ctx, ctx->bloom_lines, uint64_t(a.bh) * 16)))
		return r;
	a.buf0 = static_cast<uint64_t*>(ctx->bloom0.p);
	a.buf1 = static_cast<uint64_t*>(ctx->bloom1.p);
	a.line_sums = static_cast<int32_t*>(ctx->bloom_lines.p);
	a.average = reinterpret_cast<float*>(ctx->counters_host_dev + HOST_POST_OFFSET + 4);
	{
		StageTimer t(ctx, STAGE_BLOOM);
		swk::launch_bloom(ctx->stream, a, out_average_rgb != nullptr);
	}
	if((r = hip_check(ctx, hipGetLastError(), "bloom launch")))
		return r;
	if(out_average_rgb)
	{
		// the average kernel stores it straight into the pinned host words
		if((r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")))
			return r;
		volatile const float* avg = reinterpret_cast<volatile const float*>(ctx->counters_host + HOST_POST_OFFSET + 4);
		// lanes B, G, R -> RGB
		out_average_rgb[0] = avg[2];
		out_average_rgb[1] = avg[1];
		out_average_rgb[2] = avg[0];
	}
	return SW_OK;
}

// Small launch arguments travel from pinned staging memory. stage_begin takes the next slot of the
// ring, waits until the copies out of it (ARG_STAGING_SLOTS calls ago) have completed (its event),
// grows it if needed and returns it; stage_copy queues a copy out of the slot (at src_offset) and
// records the slot's event after it.
static int stage_begin(sw_ctx* ctx, size_t bytes, char** out)
{
	int r;
	const uint32_t k = (ctx->arg_staging_slot + 1) % ARG_STAGING_SLOTS;
	ctx->arg_staging_slot = k;
	if(ctx->arg_staging_event[k] && (r = hip_check(ctx, hipEventSynchronize(ctx->arg_staging_event[k]), "argument staging wait")))
		return r;
	if(bytes > ctx->arg_staging_cap[k])
	{
		if(ctx->arg_staging[k])
			(void)hipHostFree(ctx->arg_staging[k]);
		ctx->arg_staging[k] = nullptr;
		ctx->arg_staging_cap[k] = 0;
		const size_t cap = std::max<size_t>(bytes * 2, 64 << 10);
		if((r = hip_check(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->arg_staging[k]), cap, hipHostMallocDefault),
				"argument staging allocation")))
			return r;
		ctx->arg_staging_cap[k] = cap;
	}
	if(!ctx->arg_staging_event[k] &&
		(r = hip_check(ctx, hipEventCreateWithFlags(&ctx->arg_staging_event[k], hipEventDisableTiming), "argument staging event")))
		return r;
	*out = ctx->arg_staging[k];
	return SW_OK;
}

static int stage_copy(sw_ctx* ctx, void* dst_device, size_t bytes, size_t src_offset = 0)
{
	int r;
	const uint32_t k = ctx->arg_staging_slot;
	if((r = hip_check(ctx, hipMemcpyAsync(dst_device, ctx->arg_staging[k] + src_offset, bytes, hipMemcpyHostToDevice, ctx->stream),
			"argument copy")))
		return r;
	return hip_check(ctx, hipEventRecord(ctx->arg_staging_event[k], ctx->stream), "argument staging event");
}

int sw_mix_portal_textures(sw_ctx* ctx, sw_map* map, uint32_t color_bits, const sw_portal_mix* mixes, uint32_t num_mixes,
	void* portal_texels_device, uint64_t num_portal_texels)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!map || (num_mixes && (!mixes || !portal_texels_device)))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(color_bits != 32 && color_bits != 64)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "color_bits must be 32 or 64");
	if(num_mixes == 0)
		return SW_OK;
	std::vector<swk::PortalMixRec> recs(num_mixes);
	std::vector<std::pair<uint64_t, uint64_t>> ranges(num_mixes);
	uint64_t max_texels = 0;
	for(uint32_t i = 0; i < num_mixes; ++i)
	{
		const sw_portal_mix& x = mixes[i];
		const std::string which = "portal mix " + std::to_string(i) + ": ";
		if(x.level >= map->host_levels.size() || !(map->host_levels[x.level].flags & SW_TEX_LITE_LEVEL))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "level is not a lite level of the map");
		const sw_texture_level& lv = map->host_levels[x.level];
		if(lv.width == 0 || lv.height == 0 || lv.offset + uint64_t(lv.width) * lv.height > map->ptrs.num_lite)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "empty or out-of-range level");
		if(x.blend > SW_BLEND_ALPHA_BLEND)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "unknown blending mode");
		const uint64_t n = uint64_t(x.size[0]) * x.size[1];
		if(x.dst_offset > num_portal_texels || n > num_portal_texels - x.dst_offset)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "rect outside the portal textures");
		ranges[i] = {x.dst_offset, x.dst_offset + n};
		max_texels = std::max(max_texels, n);
		swk::PortalMixRec& r = recs[i];
		r.dst = x.dst_offset;
		r.w = x.size[0];
		r.h = x.size[1];
		r.tcu = x.tc_start[0];
		r.tcv = x.tc_start[1];
		r.tex_offset = uint32_t(lv.offset);
		r.tw = lv.width;
		r.th = lv.height;
		r.blend = x.blend;
		r.light[0] = x.light[0];
		r.light[1] = x.light[1];
		r.light[2] = x.light[2];
		r.reserved = 0;
	}
	// one writer per texel: the rects of one call must not overlap
	std::sort(ranges.begin(), ranges.end());
	for(uint32_t i = 1; i < num_mixes; ++i)
		if(ranges[i].first < ranges[i - 1].second && ranges[i].first != ranges[i].second &&
			ranges[i - 1].first != ranges[i - 1].second)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "portal mixes overlap");
	if(max_texels == 0)
		return SW_OK;
	if(max_texels > uint64_t(0xFFFFFFFFu) * 1024u)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "portal rect too large");
	(void)hipSetDevice(ctx->device);
	int r;
	const size_t bytes = sizeof(swk::PortalMixRec) * num_mixes;
	char* staging = nullptr;
	if((r = ensure(ctx, ctx->portal_mix, bytes)) || (r = stage_begin(ctx, bytes, &staging)))
		return r;
	std::memcpy(staging, recs.data(), bytes);
	if((r = stage_copy(ctx, ctx->portal_mix.p, bytes)))
		return r;
	swk::launch_portal_mix(ctx->stream, color_bits, static_cast<const uint32_t*>(map->lite.p),
		static_cast<const swk::PortalMixRec*>(ctx->portal_mix.p), num_mixes, max_texels, portal_texels_device);
	return hip_check(ctx, hipGetLastError(), "portal mix launch");
}

int sw_expand_mesh_runs(sw_ctx* ctx, const sw_mesh_run* runs, uint32_t num_runs, const sw_model_vertex* prepared_vertices_device,
	uint64_t vertex_capacity, const uint16_t* prepared_triangles_device, uint64_t triangle_capacity, const uint32_t* counts_device,
	uint32_t num_counts, int32_t* out_vertex_ints_device, uint64_t vertex_int_capacity, uint8_t* out_fan_counts)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(num_runs == 0)
		return SW_OK;
	if(!runs || !prepared_vertices_device || !prepared_triangles_device || !counts_device || !out_vertex_ints_device ||
		!out_fan_counts)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	// per run: its vertex ints and fan-count bytes, back to back; every index range checked here
	std::vector<uint64_t> bases(size_t(num_runs) * 2);
	uint64_t ints = 0, slots = 0;
	uint32_t max_triangles = 0;
	for(uint32_t i = 0; i < num_runs; ++i)
	{
		const sw_mesh_run& r = runs[i];
		const std::string which = "mesh run " + std::to_string(i) + ": ";
		if(r.count_index >= num_counts)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "count index out of range");
		if(r.num_planes3d > SW_MESH_MAX_PLANES3D || r.num_planes2d > 4)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "too many clip planes");
		if(r.max_triangles > SW_MODEL_MAX_TRIANGLES || r.num_vertices == 0 || r.num_vertices > 65536)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "mesh size out of range");
		if(uint64_t(r.first_vertex) + r.num_vertices > vertex_capacity ||
			uint64_t(r.first_triangle) + r.max_triangles > triangle_capacity)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "prepared vertices or triangles beyond the buffers");
		if(r.blend > SW_BLEND_ALPHA_BLEND)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "unknown blending mode");
		const uint64_t fan = 1u + r.num_planes3d + r.num_planes2d;
		bases[size_t(i) * 2] = ints;
		bases[size_t(i) * 2 + 1] = slots;
		ints += uint64_t(r.max_triangles) * fan * 21u;
		slots += r.max_triangles;
		max_triangles = std::max(max_triangles, r.max_triangles);
	}
	if(ints > vertex_int_capacity)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "vertex int capacity below the runs' fan slots");
	if(slots == 0)
		return SW_OK;
	(void)hipSetDevice(ctx->device);
	const size_t run_bytes = sizeof(sw_mesh_run) * num_runs, base_at = (run_bytes + 255) & ~size_t(255),
				 total = base_at + sizeof(uint64_t) * bases.size();
	int r;
	if((r = ensure(ctx, ctx->mesh_args, total)) || (r = ensure(ctx, ctx->mesh_fans, slots)))
		return r;
	char* staging = nullptr;
	if((r = stage_begin(ctx, total, &staging)))
		return r;
	std::memcpy(staging, runs, run_bytes);
	std::memcpy(staging + base_at, bases.data(), sizeof(uint64_t) * bases.size());
	if((r = stage_copy(ctx, ctx->mesh_args.p, total)))
		return r;
	{
		StageTimer t(ctx, STAGE_MODELS);
		swk::launch_expand_mesh_runs(ctx->stream, static_cast<const sw_mesh_run*>(ctx->mesh_args.p), num_runs, max_triangles,
			reinterpret_cast<const uint64_t*>(static_cast<const char*>(ctx->mesh_args.p) + base_at), prepared_vertices_device,
			prepared_triangles_device, counts_device, out_vertex_ints_device, static_cast<uint8_t*>(ctx->mesh_fans.p));
	}
	if((r = hip_check(ctx, hipGetLastError(), "mesh expansion launch")))
		return r;
	if((r = hip_check(ctx, hipMemcpyAsync(out_fan_counts, ctx->mesh_fans.p, slots, hipMemcpyDeviceToHost, ctx->stream),
			"fan counts copy")))
		return r;
	return hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int sw_prepare_models(sw_ctx* ctx, const sw_model_mesh* meshes, uint32_t num_meshes, const sw_model_instance* instances,
	uint32_t num_instances, sw_model_vertex* out_vertices, uint64_t vertex_capacity, uint16_t* out_triangles,
	uint64_t triangle_capacity, uint32_t* out_counts)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(num_instances == 0)
		return SW_OK;
	if(!meshes || !instances || !out_vertices || !out_triangles || !out_counts)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(reinterpret_cast<uintptr_t>(out_vertices) % 16)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "out_vertices must be 16-byte aligned");
	for(uint32_t i = 0; i < num_meshes; ++i)
	{
		const sw_model_mesh& m = meshes[i];
		const std::string which = "model mesh " + std::to_string(i) + ": ";
		if(m.kind > SW_MODEL_SKELETON_ANIMATED)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "unknown kind");
		if(m.num_triangles > SW_MODEL_MAX_TRIANGLES || m.num_vertices > 65536)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "more than 65536 triangles or vertices");
		if((m.num_triangles && (!m.triangles || m.num_vertices == 0)) || (m.num_vertices && !m.vertices))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "missing vertex or triangle data");
		if(m.kind == SW_MODEL_VERTEX_ANIMATED && m.num_vertices && (m.num_frames == 0 || !m.frame_vertices))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "vertex animation without frames");
		if(m.kind == SW_MODEL_SKELETON_ANIMATED &&
			(m.num_bones > SW_MODEL_MAX_BONES || (m.num_bones && (m.num_frames == 0 || !m.frame_bones || !m.bone_parents))))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "bad skeleton (more than 255 bones, or no frames)");
	}
	// per-instance checks, output ranges, and the scratch regions of sorts too large for LDS
	const uint32_t lds_limit = swk::model_sort_lds_limit();
	std::vector<uint64_t> sort_offsets(num_instances, 0);
	uint64_t scratch_entries = 0;
	for(uint32_t i = 0; i < num_instances; ++i)
	{
		const sw_model_instance& in = instances[i];
		const std::string which = "model instance " + std::to_string(i) + ": ";
		if(in.mesh >= num_meshes)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "mesh index out of range");
		const sw_model_mesh& m = meshes[in.mesh];
		const bool animated = m.kind == SW_MODEL_VERTEX_ANIMATED || (m.kind == SW_MODEL_SKELETON_ANIMATED && m.num_bones);
		if(animated && m.num_vertices && (in.frames[0] >= m.num_frames || in.frames[1] >= m.num_frames))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "animation frame out of range");
		if(uint64_t(in.first_vertex) + m.num_vertices > vertex_capacity ||
			uint64_t(in.first_triangle) + m.num_triangles > triangle_capacity)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, which + "output range beyond the buffers' capacity");
		uint64_t n_pad = 1;
		while(n_pad < m.num_triangles)
			n_pad <<= 1;
		if(n_pad > lds_limit)
		{
			sort_offsets[i] = scratch_entries;
			scratch_entries += n_pad;
		}
	}
	(void)hipSetDevice(ctx->device);
	const size_t mesh_bytes = sizeof(sw_model_mesh) * num_meshes, inst_bytes = sizeof(sw_model_instance) * num_instances,
				 off_bytes = sizeof(uint64_t) * num_instances;
	const size_t inst_at = (mesh_bytes + 255) & ~size_t(255), off_at = (inst_at + inst_bytes + 255) & ~size_t(255);
	int r;
	if((r = ensure(ctx, ctx->model_args, off_at + off_bytes)) ||
		(scratch_entries && (r = ensure(ctx, ctx->model_scratch, scratch_entries * sizeof(uint64_t)))))
		return r;
	char* args = static_cast<char*>(ctx->model_args.p);
	char* staging = nullptr;
	if((r = stage_begin(ctx, off_at + off_bytes, &staging)))
		return r;
	std::memcpy(staging, meshes, mesh_bytes);
	std::memcpy(staging + inst_at, instances, inst_bytes);
	std::memcpy(staging + off_at, sort_offsets.data(), off_bytes);
	if((r = stage_copy(ctx, args, off_at + off_bytes)))
		return r;
	{
		StageTimer t(ctx, STAGE_MODELS);
		swk::launch_prepare_models(ctx->stream, num_instances, reinterpret_cast<const sw_model_mesh*>(args),
			reinterpret_cast<const sw_model_instance*>(args + inst_at), reinterpret_cast<const uint64_t*>(args + off_at),
			static_cast<uint64_t*>(ctx->model_scratch.p), out_vertices, out_triangles, out_counts,
			ctx->counters_host_dev + HOST_MODEL_ERR_OFFSET);
	}
	return hip_check(ctx, hipGetLastError(), "model preparation launch");
}

int sw_apply_texture_ops(
	sw_ctx* ctx, sw_map* map, const sw_texture_op* ops, uint32_t num_ops, const int32_t* turb_shifts, uint32_t num_turb_shifts)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!map || (num_ops && !ops) || (num_turb_shifts && !turb_shifts))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(num_ops == 0)
		return SW_OK;
	swk::TexOpPlan plan;
	std::string error;
	if(swk::pack_texture_ops(ops, num_ops, map->host_levels.data(), uint32_t(map->host_levels.size()), num_turb_shifts, ctx->texop_packed,
		   plan, error))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, error);
	(void)hipSetDevice(ctx->device);
	const size_t shifts_at = (plan.bytes + 255) & ~size_t(255), total = shifts_at + sizeof(int32_t) * num_turb_shifts;
	int r;
	char* staging = nullptr;
	if((r = ensure(ctx, ctx->texop_args, total)) || (r = stage_begin(ctx, total, &staging)))
		return r;
	std::memcpy(staging, ctx->texop_packed.data(), plan.bytes);
	if(num_turb_shifts)
		std::memcpy(staging + shifts_at, turb_shifts, sizeof(int32_t) * num_turb_shifts);
	char* args = static_cast<char*>(ctx->texop_args.p);
	if((r = stage_copy(ctx, args, total)))
		return r;
	{
		StageTimer t(ctx, STAGE_TEXTURES);
		swk::launch_texture_ops(ctx->stream, args, plan, reinterpret_cast<const int32_t*>(args + shifts_at),
			static_cast<sw_texel*>(map->texels.p), static_cast<uint32_t*>(map->lite.p));
	}
	return hip_check(ctx, hipGetLastError(), "texture ops launch");
}

static bool water_size_ok(uint32_t log2w, uint32_t log2h)
{
	// GenerativeTextureEffectWater::new (generative_texture_effect_water.rs:21-34): sides of at least
	// 2^MAX_MIP, area below 2^22
	return log2w >= 4 && log2h >= 4 && log2w + log2h < 22;
}

int sw_water_step(sw_ctx* ctx, float* field_device, float* field_old_device, uint32_t log2_width, uint32_t log2_height,
	float attenuation, const sw_wave_point* points, uint32_t num_points)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!field_device || !field_old_device || (num_points && !points))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(!water_size_ok(log2_width, log2_height))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water field size out of range (sides >= 16, area < 2^22)");
	if(field_device == field_old_device)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "the two field buffers must differ");
	(void)hipSetDevice(ctx->device);
	int r;
	const sw_wave_point* pts_dev = nullptr;
	if(num_points)
	{
		const size_t bytes = sizeof(sw_wave_point) * num_points;
		char* staging = nullptr;
		if((r = ensure(ctx, ctx->water_points, bytes)) || (r = stage_begin(ctx, bytes, &staging)))
			return r;
		std::memcpy(staging, points, bytes);
		if((r = stage_copy(ctx, ctx->water_points.p, bytes)))
			return r;
		pts_dev = static_cast<const sw_wave_point*>(ctx->water_points.p);
	}
	{
		StageTimer t(ctx, STAGE_TEXTURES);
		swk::launch_water_step(ctx->stream, field_device, field_old_device, log2_width, log2_height, attenuation, pts_dev, num_points);
	}
	return hip_check(ctx, hipGetLastError(), "water step launch");
}

int sw_fire_step(sw_ctx* ctx, uint8_t* heat_device, uint8_t* heat_old_device, uint32_t log2_width, uint32_t log2_height,
	float attenuation, uint32_t slow, const sw_heat_point* points, uint32_t num_points)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!heat_device || !heat_old_device || (num_points && !points))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(!water_size_ok(log2_width, log2_height))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "heat map size out of range (sides >= 16, area < 2^22)");
	if(heat_device == heat_old_device)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "the two heat buffers must differ");
	// update_heat_map (generative_texture_effect_fire.rs:359): attenuation in 2^-18 steps
	const uint32_t attenuation_i = uint32_t(std::fmin(std::fmax(attenuation, 0.0f), 1.0f) * float(1u << 18));
	(void)hipSetDevice(ctx->device);
	int r;
	const sw_heat_point* pts_dev = nullptr;
	if(num_points)
	{
		const size_t bytes = sizeof(sw_heat_point) * num_points;
		char* staging = nullptr;
		if((r = ensure(ctx, ctx->fire_args, bytes)) || (r = stage_begin(ctx, bytes, &staging)))
			return r;
		std::memcpy(staging, points, bytes);
		if((r = stage_copy(ctx, ctx->fire_args.p, bytes)))
			return r;
		pts_dev = static_cast<const sw_heat_point*>(ctx->fire_args.p);
	}
	{
		StageTimer t(ctx, STAGE_TEXTURES);
		swk::launch_fire_step(ctx->stream, heat_device, heat_old_device, log2_width, log2_height, attenuation_i, slow != 0, pts_dev, num_points);
	}
	return hip_check(ctx, hipGetLastError(), "fire step launch");
}

int sw_fire_make_texture(sw_ctx* ctx, sw_map* map, const uint8_t* heat_device, uint32_t log2_width, uint32_t log2_height,
	uint32_t dst_level, const uint32_t palette[256])
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!map || !heat_device || !palette)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(!water_size_ok(log2_width, log2_height))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "heat map size out of range (sides >= 16, area < 2^22)");
	const std::vector<sw_texture_level>& lv = map->host_levels;
	if(uint64_t(dst_level) + 4 >= lv.size())
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "fire destination level (and its 4 mips) out of the level table");
	uint64_t off[5];
	for(uint32_t mip = 0; mip <= 4; ++mip)
	{
		const sw_texture_level& l = lv[dst_level + mip];
		if(l.width != (1u << (log2_width - mip)) || l.height != (1u << (log2_height - mip)) || !(l.flags & SW_TEX_LITE_LEVEL))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "fire destination levels must be lite levels of the heat map's size and its mips");
		off[mip] = l.offset;
	}
	(void)hipSetDevice(ctx->device);
	int r;
	char* staging = nullptr;
	if((r = ensure(ctx, ctx->fire_args, 1024)) || (r = stage_begin(ctx, 1024, &staging)))
		return r;
	std::memcpy(staging, palette, 1024);
	if((r = stage_copy(ctx, ctx->fire_args.p, 1024)))
		return r;
	{
		StageTimer t(ctx, STAGE_TEXTURES);
		swk::launch_fire_texture(ctx->stream, heat_device, log2_width, log2_height, off, static_cast<const uint32_t*>(ctx->fire_args.p),
			static_cast<uint32_t*>(map->lite.p));
	}
	return hip_check(ctx, hipGetLastError(), "fire texture launch");
}

int sw_water_make_texture(sw_ctx* ctx, sw_map* map, const float* field_device, uint32_t log2_width, uint32_t log2_height,
	const sw_water_texture_params* params)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!map || !field_device || !params)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(!water_size_ok(log2_width, log2_height))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water field size out of range (sides >= 16, area < 2^22)");
	const sw_water_texture_params& p = *params;
	if(p.color_mode > SW_WATER_COLOR_DEFORMED_X || p.init > SW_WATER_INIT_SOURCE)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "unknown water colour mode or init");
	const std::vector<sw_texture_level>& lv = map->host_levels;
	if(uint64_t(p.dst_level) + 4 >= lv.size())
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water destination level (and its 4 mips) out of the level table");
	swk::WaterTexLaunch w;
	for(uint32_t mip = 0; mip <= 4; ++mip)
	{
		const sw_texture_level& l = lv[p.dst_level + mip];
		if(l.width != (1u << (log2_width - mip)) || l.height != (1u << (log2_height - mip)) || (l.flags & SW_TEX_LITE_LEVEL))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water destination levels must be the field's size and its mips");
		w.dst_offset[mip] = l.offset;
	}
	if(!(lv[p.dst_level].flags & SW_TEX_HAS_NORMAL_MAP))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water destination level lacks SW_TEX_HAS_NORMAL_MAP");
	const bool needs_color = p.color_mode != SW_WATER_COLOR_KEEP || p.init == SW_WATER_INIT_SOURCE;
	w.color_offset = 0;
	if(needs_color)
	{
		if(p.color_level >= lv.size())
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water colour level out of the level table");
		const sw_texture_level& c = lv[p.color_level];
		if(c.width != (1u << log2_width) || c.height != (1u << log2_height) || (c.flags & SW_TEX_LITE_LEVEL))
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water colour level must have the field's size");
		if(p.color_level >= p.dst_level && p.color_level <= p.dst_level + 4)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "water colour level is one of the destinations");
		w.color_offset = c.offset;
	}
	w.color_mode = p.color_mode;
	w.init = p.init;
	// texture_data.material.roughness.max(MIN_VALID_ROUGHNESS).min(1.0) (:283)
	w.roughness = std::fmin(std::fmax(p.roughness, 1.0f / 126.0f), 1.0f);
	w.base_color = p.base_color;
	(void)hipSetDevice(ctx->device);
	{
		StageTimer t(ctx, STAGE_TEXTURES);
		swk::launch_water_texture(ctx->stream, field_device, log2_width, log2_height, w, static_cast<sw_texel*>(map->texels.p));
	}
	return hip_check(ctx, hipGetLastError(), "water texture launch");
}

int sw_render_depth(sw_ctx* ctx, const sw_depth_target* targets, uint32_t num_targets, const sw_depth_draw* draws, uint32_t num_draws,
	const int32_t* vertex_data, uint64_t num_vertex_ints, float* depth_texels_device, uint64_t num_depth_texels)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if((num_targets && !targets) || (num_draws && (!draws || !vertex_data)) || !depth_texels_device)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(num_targets == 0)
		return SW_OK;
	swk::DepthPlan plan;
	std::string error;
	if(swk::depth_plan(targets, num_targets, draws, num_draws, num_vertex_ints, num_depth_texels, ctx->depth_packed, plan, error))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, error);
	(void)hipSetDevice(ctx->device);
	int r;
	const size_t vbytes = sizeof(int32_t) * num_vertex_ints;
	if((r = ensure(ctx, ctx->depth_args, plan.bytes)) || (r = ensure(ctx, ctx->depth_vdata, std::max<size_t>(vbytes, 4))) ||
		(r = ensure(ctx, ctx->depth_rows, std::max<uint64_t>(plan.row_records, 1) * swk::depth_row_record_bytes())) ||
		(r = ensure(ctx, ctx->depth_draw_rows, sizeof(int32_t) * 2 * std::max<uint32_t>(num_draws, 1))))
		return r;
	// the packed draws and the vertex data through one staging slot (no copy from pageable memory)
	char* staging = nullptr;
	const size_t vat = (plan.bytes + 255) & ~size_t(255);
	if((r = stage_begin(ctx, vat + vbytes, &staging)))
		return r;
	std::memcpy(staging, ctx->depth_packed.data(), plan.bytes);
	if(vbytes)
		std::memcpy(staging + vat, vertex_data, vbytes);
	if((r = stage_copy(ctx, ctx->depth_args.p, plan.bytes)) || (vbytes && (r = stage_copy(ctx, ctx->depth_vdata.p, vbytes, vat))))
		return r;
	{
		StageTimer t(ctx, STAGE_DEPTH);
		swk::launch_depth(ctx->stream, static_cast<const char*>(ctx->depth_args.p), plan, static_cast<const int32_t*>(ctx->depth_vdata.p),
			ctx->depth_rows.p, static_cast<int32_t*>(ctx->depth_draw_rows.p), depth_texels_device);
	}
	return hip_check(ctx, hipGetLastError(), "depth launch");
}

int sw_set_light_shadows(
	sw_ctx* ctx, const sw_light_shadow* shadows, uint32_t num, const float* shadow_texels_device, uint64_t num_shadow_texels)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(num == 0)
	{
		if(ctx->num_shadows)
			ctx->alloc_epoch++;
		ctx->num_shadows = 0;
		ctx->shadow_texels = nullptr;
		return SW_OK;
	}
	if(!shadows || !shadow_texels_device)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	for(uint32_t i = 0; i < num; ++i)
	{
		const sw_light_shadow& sh = shadows[i];
		if(sh.kind > SW_SHADOW_PROJECTOR)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "light shadow " + std::to_string(i) + ": unknown kind");
		if(sh.kind == SW_SHADOW_NONE)
			continue;
		const uint64_t n = uint64_t(sh.size) * sh.size * (sh.kind == SW_SHADOW_CUBE ? 6u : 1u);
		if(sh.size == 0 || sh.size > 65536 || sh.offset > num_shadow_texels || n > num_shadow_texels - sh.offset)
			return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "light shadow " + std::to_string(i) + ": map outside the shadow texels");
	}
	(void)hipSetDevice(ctx->device);
	int r;
	if((r = ensure(ctx, ctx->shadows, sizeof(sw_light_shadow) * num)))
		return r;
	char* staging = nullptr;
	if((r = stage_begin(ctx, sizeof(sw_light_shadow) * num, &staging)))
		return r;
	std::memcpy(staging, shadows, sizeof(sw_light_shadow) * num);
	if((r = stage_copy(ctx, ctx->shadows.p, sizeof(sw_light_shadow) * num)))
		return r;
	ctx->num_shadows = num;
	ctx->shadow_texels = shadow_texels_device;
	ctx->alloc_epoch++; // a captured graph holds the previous shadow arguments
	return SW_OK;
}

int sw_make_skybox_side(sw_ctx* ctx, sw_map* map, uint32_t color_bits, const void* image_device, uint32_t image_width,
	uint32_t image_height, float brightness, uint64_t dst_offset, uint32_t* out_size)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!map || !image_device || (color_bits != 32 && color_bits != 64))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument or colour size not 32 / 64");
	const uint32_t size = std::min(image_width, image_height);
	if(size == 0 || size > 65536)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "skybox image size out of range");
	uint64_t total = 0;
	for(uint32_t mip = 0, s = size; mip < 5 && s; ++mip, s >>= 1)
		total += uint64_t(s) * s;
	const uint64_t cap = color_bits == 64 ? map->num_sky64 : map->num_sky32;
	if(dst_offset > cap || total > cap - dst_offset)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "skybox side beyond the map's skybox buffer");
	// (brightness * 256) as i32: truncation, saturating, NaN -> 0
	const float bf = brightness * 256.0f;
	const int32_t bi = bf != bf ? 0 : (bf <= -2147483648.0f ? INT32_MIN : (bf >= 2147483648.0f ? INT32_MAX : int32_t(bf)));
	(void)hipSetDevice(ctx->device);
	char* dst = static_cast<char*>(color_bits == 64 ? map->sky64.p : map->sky32.p) + dst_offset * (color_bits / 8);
	swk::launch_skybox_side(ctx->stream, color_bits, image_device, image_width, size, bi, dst);
	if(out_size)
		*out_size = size;
	return hip_check(ctx, hipGetLastError(), "skybox launch");
}

int sw_map_read(sw_ctx* ctx, const sw_map* map, uint32_t which, uint64_t first, uint64_t count, void* dst)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!map || (count && !dst) || which > 3)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument or unknown buffer");
	const uint64_t ns[4] = {map->ptrs.num_texels, map->ptrs.num_lite, map->num_sky32, map->num_sky64};
	const size_t es[4] = {sizeof(sw_texel), sizeof(uint32_t), sizeof(uint32_t), sizeof(uint64_t)};
	const void* bufs[4] = {map->texels.p, map->lite.p, map->sky32.p, map->sky64.p};
	const uint64_t n = ns[which];
	const size_t elem = es[which];
	if(first > n || count > n - first)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "range beyond the map's texels");
	if(count == 0)
		return SW_OK;
	(void)hipSetDevice(ctx->device);
	const char* src = static_cast<const char*>(bufs[which]) + first * elem;
	int r;
	if((r = hip_check(ctx, hipMemcpyAsync(dst, src, count * elem, hipMemcpyDeviceToHost, ctx->stream), "map read")))
		return r;
	return hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

float sw_update_exposure(const sw_exposure_params* p, float current_exposure, const float average_rgb[3], float frame_duration_s)
{
	if(!p || !average_rgb)
		return current_exposure;
	// synchronize_config (postprocessor.rs:928-940) clamps
	const float min_exposure = std::min(std::max(p->min_exposure, 1.0f / 65536.0f), 16.0f);
	const float max_exposure = std::min(std::max(p->max_exposure, 1.0f), 65536.0f);
	const float speed = std::min(std::max(p->exposure_update_speed, 0.5f), 16.0f);
	const float base = std::min(std::max(p->base_brightness, 0.5f), 2.0f);
	const float zero_level = std::min(std::max(p->zero_level_brightness, 1.0f / 16.0f), 16.0f);
	const float power = std::min(std::max(p->brightness_scale_power, 0.01f), 0.5f);
	// get_color_brightness (:1047-1051), update_exposure (:206-223)
	const float brightness =
		std::min(std::max((average_rgb[0] + average_rgb[1] + average_rgb[2]) * 0.33f, 1.0f / 1024.0f), 65536.0f) / 255.0f;
	const float target = std::pow(brightness / zero_level, power) * 0.5f * base / brightness;
	const float clamped = std::min(std::max(target, min_exposure), max_exposure);
	const float mix = std::exp(-speed * frame_duration_s);
	return std::exp(std::log(current_exposure) * mix + std::log(clamped) * (1.0f - mix));
}

void* sw_frame_device_framebuffer(sw_ctx* ctx) { return ctx ? ctx->fb.p : nullptr; }

int sw_device_alloc(sw_ctx* ctx, uint64_t bytes, void** out)
{
	if(!ctx || !out)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	(void)hipSetDevice(ctx->device);
	*out = nullptr;
	if(hipMalloc(out, std::max<uint64_t>(bytes, 16)) != hipSuccess)
	{
		(void)hipGetLastError();
		*out = nullptr;
		return set_error(ctx, SW_ERR_OUT_OF_MEMORY, "hipMalloc failed");
	}
	return SW_OK;
}

void sw_device_free(sw_ctx* ctx, void* p)
{
	if(!ctx || !p)
		return;
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	(void)hipFree(p);
}

int sw_copy_to_device(sw_ctx* ctx, void* dst, const void* src, uint64_t bytes)
{
	if(!ctx || ((!dst || !src) && bytes))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	(void)hipSetDevice(ctx->device);
	if(!bytes)
		return SW_OK;
	int r = hip_check(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), "copy to device");
	return r ? r : hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int sw_copy_from_device(sw_ctx* ctx, void* dst, const void* src, uint64_t bytes)
{
	if(!ctx || ((!dst || !src) && bytes))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	(void)hipSetDevice(ctx->device);
	if(!bytes)
		return SW_OK;
	int r = hip_check(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), "copy from device");
	return r ? r : hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int sw_copy_to_device_async(sw_ctx* ctx, void* dst, const void* src, uint64_t bytes)
{
	if(!ctx || ((!dst || !src) && bytes))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	(void)hipSetDevice(ctx->device);
	if(!bytes)
		return SW_OK;
	return hip_check(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), "copy to device");
}

int sw_copy_from_device_async(sw_ctx* ctx, void* dst, const void* src, uint64_t bytes)
{
	if(!ctx || ((!dst || !src) && bytes))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	(void)hipSetDevice(ctx->device);
	if(!bytes)
		return SW_OK;
	return hip_check(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), "copy from device");
}

int sw_host_alloc(sw_ctx* ctx, uint64_t bytes, void** out)
{
	if(!ctx || !out)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	(void)hipSetDevice(ctx->device);
	*out = nullptr;
	if(hipHostMalloc(out, std::max<uint64_t>(bytes, 16), hipHostMallocDefault) != hipSuccess)
	{
		(void)hipGetLastError();
		*out = nullptr;
		return set_error(ctx, SW_ERR_OUT_OF_MEMORY, "hipHostMalloc failed");
	}
	return SW_OK;
}

void sw_host_free(sw_ctx* ctx, void* p)
{
	if(!ctx || !p)
		return;
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	(void)hipHostFree(p);
}

int sw_profile_enable(sw_ctx* ctx, uint32_t enable)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	ctx->profiling = enable != 0;
	return SW_OK;
}

int sw_profile_read(sw_ctx* ctx, double out_ms[SW_PROFILE_STAGES], uint64_t out_counts[SW_PROFILE_STAGES], uint32_t reset)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	int r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
	if(r)
		return r;
	collect_timings(ctx);
	for(int i = 0; i < SW_PROFILE_STAGES; ++i)
	{
		if(out_ms)
			out_ms[i] = ctx->prof_ms[i];
		if(out_counts)
			out_counts[i] = ctx->prof_count[i];
		if(reset)
		{
			ctx->prof_ms[i] = 0.0;
			ctx->prof_count[i] = 0;
		}
	}
	return SW_OK;
}

int sw_read_surfaces(sw_ctx* ctx, void* dst, uint64_t num_texels)
{
	if(!ctx || (!dst && num_texels))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	(void)hipSetDevice(ctx->device);
	int r;
	if(ctx->frame_active && ctx->frame_surfaces_pending && (r = launch_frame(ctx, 0, 0, false)))
		return r;
	const uint64_t bytes = num_texels * color_bytes(ctx->last_color_bits);
	if(bytes > ctx->arena.cap)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "num_texels beyond the surface arena");
	if(bytes == 0)
		return SW_OK;
	r = hip_check(ctx, hipMemcpyAsync(dst, ctx->arena.p, bytes, hipMemcpyDeviceToHost, ctx->stream), "read surfaces");
	return r ? r : hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int sw_get_stats(sw_ctx* ctx, sw_stats* out)
{
	if(!ctx || !out)
		return SW_ERR_INVALID_ARGUMENT;
	(void)hipSetDevice(ctx->device);
	int r = ensure_stats(ctx);
	if(r)
		return r;
	*out = ctx->stats;
	return SW_OK;
}

uint32_t sw_last_reruns(const sw_ctx* ctx) { return ctx ? uint32_t(ctx->stats.reruns) : 0u; }

// ---------------------------------------------------------------------------------------------
// Batched path
// ---------------------------------------------------------------------------------------------
// Validates a batch, makes its views current (inline for one view unless `for_graph`: a captured
// graph reads the views from the device array, which every replay refreshes) and fills L.
static int prepare_batch(sw_ctx* ctx, sw_map* map, const sw_batch* b, void* out_fb, sw_ctx::Launch& L, std::vector<ViewExt>& ve,
	bool for_graph = false)
{
	if(!ctx || !map || !b || !out_fb || (b->num_views && !b->views))
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null argument");
	if(b->color_bits != 32 && b->color_bits != 64)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "color_bits must be 32 or 64");
	int r;
	if((r = build_view_ext(ctx, b->views, b->num_views, b->surface_arena_texels, ve, L.total_tiles, &L.max_view_height)))
		return r;
	if((r = validate_views_fit(ctx, ve, b->framebuffer_texels, b->num_surfaces, b->num_draws)))
		return r;
	if((r = set_views(ctx, ve, !for_graph)))
		return r;
	if((r = ensure(ctx, ctx->arena, std::max<uint64_t>(b->surface_arena_texels, 1) * color_bytes(b->color_bits))))
		return r;
	L.map = map;
	L.color_bits = b->color_bits;
	L.num_views = b->num_views;
	L.num_surfaces = b->num_surfaces;
	L.num_lights = b->lights ? b->num_lights : 0;
	L.num_light_indices = b->light_indices ? b->num_light_indices : 0;
	L.num_draws = b->num_draws;
	L.num_vertex_ints = b->num_vertex_ints;
	L.arena_texels = b->surface_arena_texels;
	L.portal = b->portal_texels;
	L.num_portal = b->portal_texels ? b->num_portal_texels : 0;
	L.arena = ctx->arena.p;
	L.fb = out_fb;
	if(b->inputs_on_device)
	{
		L.surfaces = b->surfaces;
		L.lights = b->lights;
		L.light_indices = b->light_indices;
		L.draws = b->draws;
		L.vdata = b->vertex_data;
	}
	else
	{
		if((r = upload(ctx, ctx->surfaces, b->surfaces, sizeof(sw_surface_desc) * b->num_surfaces)))
			return r;
		if((r = upload(ctx, ctx->lights, b->lights, sizeof(sw_light) * L.num_lights)))
			return r;
		if((r = upload(ctx, ctx->light_indices, b->light_indices, sizeof(uint32_t) * L.num_light_indices)))
			return r;
		if((r = upload(ctx, ctx->draws, b->draws, sizeof(sw_draw) * b->num_draws)))
			return r;
		if((r = upload(ctx, ctx->vdata, b->vertex_data, sizeof(int32_t) * b->num_vertex_ints)))
			return r;
		L.surfaces = static_cast<const sw_surface_desc*>(ctx->surfaces.p);
		L.lights = static_cast<const sw_light*>(ctx->lights.p);
		L.light_indices = static_cast<const uint32_t*>(ctx->light_indices.p);
		L.draws = static_cast<const sw_draw*>(ctx->draws.p);
		L.vdata = static_cast<const int32_t*>(ctx->vdata.p);
		fill_surface_stats(ctx, b->surfaces, b->num_surfaces, true);
		if(b->num_views > 1)
		{
			if((r = host_prepare_surfaces(ctx, map, ve, b->surfaces, b->num_surfaces, b->surface_arena_texels)))
				return r;
			L.surface_mode = SURF_TABLE_HOST;
		}
	}
	if(b->num_views == 1)
		L.surface_mode = SURF_SEARCH;
	return SW_OK;
}

int sw_render_batch(sw_ctx* ctx, sw_map* map, const sw_batch* b, void* out_fb)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	(void)hipSetDevice(ctx->device);
	ctx->graph_key = 0; // the views / command buffers a captured graph reads are about to change
	std::vector<ViewExt> ve;
	sw_ctx::Launch L;
	int r;
	if((r = prepare_batch(ctx, map, b, out_fb, L, ve)))
		return r;
	ctx->last = L;
	ctx->last_color_bits = L.color_bits;
	// views that ask for a cleared background get it in the raster kernel's tile init
	if((r = launch_stages(ctx, L, true, true)))
		return r;
	return finish(ctx);
}

int sw_render_batch_async(sw_ctx* ctx, sw_map* map, const sw_batch* b, void* out_fb)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(ctx->batch_in_flight)
		return set_error(ctx, SW_ERR_STATE, "a batch is already in flight on this context");
	(void)hipSetDevice(ctx->device);
	ctx->graph_key = 0;
	std::vector<ViewExt> ve;
	sw_ctx::Launch L;
	int r;
	if((r = prepare_batch(ctx, map, b, out_fb, L, ve)))
		return r;
	ctx->last = L;
	ctx->last_color_bits = L.color_bits;
	if((r = launch_stages(ctx, L, true, true)))
		return r;
	ctx->batch_in_flight = true;
	return SW_OK;
}

int sw_batch_poll(sw_ctx* ctx)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!ctx->batch_in_flight)
		return 1;
	const hipError_t e = hipStreamQuery(ctx->stream);
	if(e == hipSuccess)
		return 1;
	if(e == hipErrorNotReady)
		return 0;
	return hip_check(ctx, e, "hipStreamQuery");
}

int sw_batch_wait(sw_ctx* ctx)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!ctx->batch_in_flight)
		return SW_OK;
	(void)hipSetDevice(ctx->device);
	ctx->batch_in_flight = false;
	return finish(ctx);
}

int sw_render_batch_graph(sw_ctx* ctx, sw_map* map, const sw_batch* b, void* out_fb)
{
	if(!ctx || !b)
		return SW_ERR_INVALID_ARGUMENT;
	if(!b->inputs_on_device)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "graph replay needs device-resident inputs");
	(void)hipSetDevice(ctx->device);
	// key: pointers and sizes that the captured launches bake in
	uint64_t key = 1469598103934665603ull;
	auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
	mix(uint64_t(uintptr_t(map)));
	mix(uint64_t(uintptr_t(out_fb)));
	mix(uint64_t(uintptr_t(b->surfaces)));
	mix(uint64_t(uintptr_t(b->draws)));
	mix(uint64_t(uintptr_t(b->vertex_data)));
	mix(b->num_views);
	mix(b->num_surfaces);
	mix(b->num_draws);
	mix(b->color_bits);
	mix(b->surface_arena_texels);
	mix(uint64_t(uintptr_t(b->portal_texels)));
	mix(b->num_portal_texels);
	mix(ctx->tile_rows);
	mix(ctx->scalar_spans ? 1 : 0);
	for(uint32_t i = 0; i < b->num_views; ++i)
	{
		mix(b->views[i].width);
		mix(b->views[i].height);
		mix(b->views[i].num_draws);
		mix(b->views[i].num_surfaces);
	}
	// content of the views (camera positions, framebuffer offsets, clear flags may change between
	// replays without a new capture): a replay with the same views skips rebuilding and uploading them
	static_assert(sizeof(sw_view) % 8 == 0, "sw_view hashed as 64-bit words");
	uint64_t views_hash = 1469598103934665603ull ^ b->framebuffer_texels;
	for(uint32_t i = 0; i < b->num_views; ++i)
	{
		uint64_t w[sizeof(sw_view) / 8];
		std::memcpy(w, &b->views[i], sizeof(w));
		for(uint64_t x : w)
			views_hash = (views_hash ^ x) * 1099511628211ull;
	}
	int r;
	if(!ctx->graph_exec || key != ctx->graph_key || ctx->graph_epoch != ctx->alloc_epoch)
	{
		// An eager run sizes every buffer (span capacity included) before capture.
		if((r = sw_render_batch(ctx, map, b, out_fb)))
			return r;
		if((r = ensure_stats(ctx)))
			return r;
		ctx->span_capacity = std::max<uint64_t>(ctx->span_capacity, ctx->stats.spans + ctx->stats.spans / 2 + 1024);
		std::vector<ViewExt> ve;
		sw_ctx::Launch L;
		if((r = prepare_batch(ctx, map, b, out_fb, L, ve, true)))
			return r;
		// allocate at the new span capacity outside the capture
		if((r = ensure(ctx, ctx->spans, sizeof(SpanRec) * ctx->span_capacity)))
			return r;
		if((r = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync")))
			return r;
		if(ctx->graph_exec)
		{
			(void)hipGraphExecDestroy(ctx->graph_exec);
			ctx->graph_exec = nullptr;
		}
		if(ctx->graph)
		{
			(void)hipGraphDestroy(ctx->graph);
			ctx->graph = nullptr;
		}
		if((r = hip_check(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture")))
			return r;
		ctx->last = L;
		ctx->capturing = true;
		const int r2 = launch_stages(ctx, L, true, true);
		hipGraph_t g = nullptr;
		const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
		ctx->capturing = false;
		if(r2)
			return r2;
		if((r = hip_check(ctx, ec, "end capture")))
			return r;
		ctx->graph = g;
		if((r = hip_check(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0), "graph instantiate")))
			return r;
		ctx->graph_key = key;
		ctx->graph_epoch = ctx->alloc_epoch;
		ctx->graph_views_hash = views_hash;
	}
	else if(views_hash != ctx->graph_views_hash)
	{
		// the captured kernels read the views from the device array: refresh it (camera
		// positions, framebuffer offsets and clear flags may change between replays)
		std::vector<ViewExt> ve;
		uint32_t tiles = 0;
		if((r = build_view_ext(ctx, b->views, b->num_views, b->surface_arena_texels, ve, tiles)) ||
			(r = validate_views_fit(ctx, ve, b->framebuffer_texels, b->num_surfaces, b->num_draws)) ||
			(r = set_views(ctx, ve, false)))
			return r;
		ctx->graph_views_hash = views_hash;
	}
	if((r = hip_check(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream), "graph launch")))
		return r;
	return finish(ctx);
}

int sw_set_tile_rows(sw_ctx* ctx, uint32_t rows)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(rows != 4 && rows != 8 && rows != 16)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "tile rows must be 4, 8 or 16");
	if(rows != ctx->tile_rows)
	{
		ctx->tile_rows = rows;
		ctx->graph_key = 0; // captured launches bake the tiling in
	}
	return SW_OK;
}

uint32_t sw_get_tile_rows(const sw_ctx* ctx) { return ctx ? ctx->tile_rows : 0; }

int sw_set_scalar_spans(sw_ctx* ctx, uint32_t enable)
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if((enable != 0) != ctx->scalar_spans)
	{
		ctx->scalar_spans = enable != 0;
		ctx->graph_key = 0;
	}
	return SW_OK;
}

uint32_t sw_get_scalar_spans(const sw_ctx* ctx) { return ctx && ctx->scalar_spans ? 1u : 0u; }

int sw_autotune_batches(sw_ctx* ctx, sw_map* map, const sw_batch* batches, uint32_t num_batches, void* out_fb, uint32_t* out_rows,
	double out_ms[SW_TILE_ROWS_COUNT])
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!batches || num_batches == 0)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "no batch to tune on");
	(void)hipSetDevice(ctx->device);
	static const uint32_t candidates[SW_TILE_ROWS_COUNT] = {4, 8, 16};
	const int ROUNDS = 5;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	int r = hip_check(ctx, hipEventCreate(&e0), "hipEventCreate");
	if(!r)
		r = hip_check(ctx, hipEventCreate(&e1), "hipEventCreate");
	double best_ms = 0.0;
	uint32_t best = ctx->tile_rows;
	bool best_scalar = ctx->scalar_spans;
	const bool profiling = ctx->profiling;
	ctx->profiling = false;
	// every tile height with both ways of loading span records, each timed over all the batches
	// (median of five rounds); out_ms keeps each height's better mode, per batch
	for(uint32_t c = 0; c < 2 * SW_TILE_ROWS_COUNT && !r; ++c)
	{
		const uint32_t rows = candidates[c / 2];
		if((r = sw_set_tile_rows(ctx, rows)))
			break;
		ctx->scalar_spans = (c & 1) != 0;
		// one untimed pass sizes the buffers
		for(uint32_t i = 0; i < num_batches && !r; ++i)
			r = sw_render_batch(ctx, map, &batches[i], out_fb);
		float t[ROUNDS] = {};
		for(int k = 0; k < ROUNDS && !r; ++k)
		{
			r = hip_check(ctx, hipEventRecord(e0, ctx->stream), "hipEventRecord");
			for(uint32_t i = 0; i < num_batches && !r; ++i)
				r = sw_render_batch(ctx, map, &batches[i], out_fb);
			if(!r)
				r = hip_check(ctx, hipEventRecord(e1, ctx->stream), "hipEventRecord");
			if(!r)
				r = hip_check(ctx, hipEventSynchronize(e1), "hipEventSynchronize");
			if(!r)
				r = hip_check(ctx, hipEventElapsedTime(&t[k], e0, e1), "hipEventElapsedTime");
		}
		if(r)
			break;
		std::sort(t, t + ROUNDS);
		const double ms = double(t[ROUNDS / 2]) / num_batches;
		if(out_ms && ((c & 1) == 0 || ms < out_ms[c / 2]))
			out_ms[c / 2] = ms;
		if(c == 0 || ms < best_ms)
		{
			best_ms = ms;
			best = rows;
			best_scalar = ctx->scalar_spans;
		}
	}
	ctx->profiling = profiling;
	if(e0)
		(void)hipEventDestroy(e0);
	if(e1)
		(void)hipEventDestroy(e1);
	const int r2 = sw_set_tile_rows(ctx, best);
	if(ctx->scalar_spans != best_scalar)
	{
		ctx->scalar_spans = best_scalar;
		ctx->graph_key = 0;
	}
	if(out_rows)
		*out_rows = best;
	return r ? r : r2;
}

int sw_autotune_batch(sw_ctx* ctx, sw_map* map, const sw_batch* b, void* out_fb, uint32_t* out_rows,
	double out_ms[SW_TILE_ROWS_COUNT])
{
	if(!ctx)
		return SW_ERR_INVALID_ARGUMENT;
	if(!b)
		return set_error(ctx, SW_ERR_INVALID_ARGUMENT, "null batch");
	return sw_autotune_batches(ctx, map, b, 1, out_fb, out_rows, out_ms);
}

} // extern "C"
