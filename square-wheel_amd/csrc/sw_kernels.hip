// CDNA4 kernels of the surface builder, span rasteriser and tonemap.
//
// A batch of views (the plain path is a batch of one view) takes four launches:
//   k_setup_draws     64 or 256 threads per draw: the edge walk of rasterizer.rs:137-271 (or the
//                     triangle split of :765-982) executed identically by the draw's threads; each
//                     part's rows are spread over them, and each writes one 32-byte span record per
//                     (part, row) holding the span's clipped x range and the start / step of its
//                     interpolants (:316-422, :500-556, :673-703, :1000-1214). Extra blocks compute
//                     every surface's arena extent, validate the surface descriptors and fill the
//                     block -> first surface table of the next kernel.
//   k_build_surfaces  one thread per surface-arena texel, flattened over all surfaces of all views:
//                     surfaces.rs:433-852 + :854-971 for that texel
//   k_raster_tiles    one wave per 64 x ROWS screen tile (ROWS = 4, 8 or 16): ballots over the
//                     view's draws pick the ones overlapping the tile, in draw order; the tile's
//                     colour lives in LDS; a part's span records for the tile rows are loaded
//                     across lanes at once and broadcast with readlane; the texel addresses of all
//                     rows are computed and fetched together, then blended in row order with
//                     write_into_framebuffer (:424-445, :558-572, :705-719, :1223-1279)
//   k_finish          one block: per-draw statistics -> result counters, live counters cleared
//   k_tonemap         one thread per pixel: postprocessor.rs:237-253
#include <algorithm>

#include "sw_device.h"
#include <hip/hip_ext.h>

#include "sw_internal.h"

using namespace swd;

// =============================================================================================
// Surfaces
// =============================================================================================

// Raises an error flag: in the live counters (for the sequence's later kernels) and in its word in
// host memory (for the host).
SW_DEV void raise_error(uint32_t* counters, uint32_t* host_err, uint32_t flag)
{
	atomicOr(&counters[COUNTER_ERROR], flag);
	host_err[__builtin_ctz(flag)] = 1u;
}

// Wave lane exchange: pull `v` from lane `src` / push `v` to lane `dst` (ds_bpermute / ds_permute).
SW_DEV int32_t lane_pull(int32_t v, uint32_t src) { return __builtin_amdgcn_ds_bpermute(int32_t(src << 2), v); }
SW_DEV int32_t lane_push(int32_t v, uint32_t dst) { return __builtin_amdgcn_ds_permute(int32_t(dst << 2), v); }
SW_DEV int32_t rl(int32_t v, uint32_t lane) { return __builtin_amdgcn_readlane(v, lane); }

// Arena extent and validity of surface s (run by the extra blocks of k_setup_draws).
SW_DEV uint32_t view_of_surface(const ViewsArg& views, uint32_t s)
{
	uint32_t lo = 0, hi = views.count;
	while(hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if(views[mid].v.first_surface <= s)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

// Whether a surface descriptor only reads inside the map's buffers and writes inside the arena.
SW_DEV bool surface_desc_ok(const sw_surface_desc& d, uint64_t end, uint64_t arena_texels, const MapPtrs& m)
{
	bool ok = end <= arena_texels && d.size[0] <= 65536u && d.size[1] <= 65536u;
	if((d.flags & SW_SURF_DIFFUSE_LAYER) &&
		(d.texture_level >= m.num_levels || (m.levels[d.texture_level].flags & SW_TEX_LITE_LEVEL)))
		ok = false;
	if((d.flags & SW_SURF_EMISSIVE) && d.emissive_level != SW_NO_TEXTURE &&
		(d.emissive_level >= m.num_levels || !(m.levels[d.emissive_level].flags & SW_TEX_LITE_LEVEL)))
		ok = false;
	if((d.flags & SW_SURF_DIFFUSE_LAYER) && d.lightmap_offset != 0 && d.size[0] > 0 && d.size[1] > 0)
	{
		// Every lightmap fetch (rows lv, lv + 1; columns lu, lu + 1) must stay inside the polygon's
		// lightmap (the reference's debug_assert!s at surfaces.rs:482, :541) and inside the buffer.
		const uint32_t sl = d.lightmap_scale_log2;
		const uint64_t num_lm = (d.flags & SW_SURF_DIRECTIONAL) ? m.num_lm_dir : m.num_lm_simple;
		if(sl > 4 || d.lightmap_size[0] < 2 || d.lightmap_size[1] < 2 ||
			((uint64_t(d.size[0]) - 1 + d.lightmap_tc_shift[0]) >> sl) + 1 >= d.lightmap_size[0] ||
			((uint64_t(d.size[1]) - 1 + d.lightmap_tc_shift[1]) >> sl) + 1 >= d.lightmap_size[1] ||
			uint64_t(d.lightmap_offset) + uint64_t(d.lightmap_size[0]) * d.lightmap_size[1] > num_lm)
			ok = false;
	}
	return ok;
}

// Arena extent, owning view and validation of surface s; also fills the block -> first surface
// table for the surface blocks whose first texel lies in [end of s - 1, end of s).
SW_DEV void prep_surface(const sw_surface_desc* __restrict__ surfaces, uint32_t s, uint32_t num_surfaces,
	const ViewsArg& views, SurfExt* __restrict__ surf_ext, uint32_t* __restrict__ blk_surf,
	uint64_t arena_texels, const MapPtrs& m, uint32_t* __restrict__ counters, uint32_t* __restrict__ host_err)
{
	const uint32_t view = view_of_surface(views, s);
	const sw_surface_desc& d = surfaces[s];
	const uint64_t start = views[view].v.surface_arena_offset + d.pixels_offset;
	const uint64_t end = start + uint64_t(d.size[0]) * d.size[1];
	surf_ext[s] = SurfExt{end, view, 0u};
	bool bad = !surface_desc_ok(d, end, views[view].arena_end, m);
	uint64_t prev_end = 0;
	if(s > 0)
	{
		// Arena ranges must follow descriptor order (the reference's bump allocator,
		// partial_renderer.rs:1617-1618): compare with the predecessor's end.
		const uint32_t pv = view_of_surface(views, s - 1);
		const sw_surface_desc& p = surfaces[s - 1];
		prev_end = views[pv].v.surface_arena_offset + p.pixels_offset + uint64_t(p.size[0]) * p.size[1];
		if(prev_end > start)
			bad = true;
	}
	if(bad)
	{
		raise_error(counters, host_err, ERROR_BAD_SURFACE);
		return;
	}
	// Blocks whose first texel lies in [prev_end, end) start their search at s; the last surface
	// also takes the blocks past its end (their texels belong to no surface).
	const uint64_t bt = surf_block_texels(arena_texels);
	const uint64_t num_blocks = (arena_texels + bt - 1) / bt;
	const uint64_t b0 = (prev_end + bt - 1) / bt;
	const uint64_t b1 = s + 1 == num_surfaces ? num_blocks : std::min<uint64_t>((end + bt - 1) / bt, num_blocks);
	for(uint64_t b = b0; b < b1; ++b)
		blk_surf[b] = s;
}

struct LmE
{
	float amb[3];
	V3 dir;
	float dev;
	float col[3];
};

// Element i of a global buffer whose byte size is below 2^32 (checked when the map is uploaded, or
// by the draw validation): the 32-bit byte offset lets the load take the buffer's base from SGPRs
// and the offset from one VGPR, with no 64-bit address arithmetic.
template <class T>
SW_DEV T elem32(const T* base, uint32_t i)
{
	typedef __attribute__((address_space(1))) const T* GlobalPtr;
	typedef __attribute__((address_space(1))) const char* GlobalBytes;
	return *(GlobalPtr)((GlobalBytes)base + i * uint32_t(sizeof(T)));
}

// Element i of the lightmaps: both kinds start with the three ambient floats, and lie in one buffer
// (simple elements at lm_simple_offset), so a 32-bit byte offset from the buffer's base selects the
// element whatever the texel's kind.
SW_DEV LmE lm_load(const MapPtrs& m, bool directional, uint32_t i)
{
	typedef __attribute__((address_space(1))) const float* GlobalFloats;
	typedef __attribute__((address_space(1))) const char* GlobalBytes;
	const uint32_t off = directional ? i * uint32_t(sizeof(sw_directional_lightmap_element)) : m.lm_simple_offset + i * 12u;
	const GlobalFloats f = (GlobalFloats)((GlobalBytes)m.lm_dir + off);
	LmE e;
	e.amb[0] = f[0];
	e.amb[1] = f[1];
	e.amb[2] = f[2];
	if(directional)
	{
		// sw_directional_lightmap_element: ambient, direction, deviation, colour
		e.dir = v3(f[3], f[4], f[5]);
		e.dev = f[6];
		e.col[0] = f[7];
		e.col[1] = f[8];
		e.col[2] = f[9];
	}
	else
	{
		e.dir = v3(0.0f, 0.0f, 0.0f);
		e.dev = 0.0f;
		e.col[0] = e.col[1] = e.col[2] = 0.0f;
	}
	return e;
}

// surfaces.rs:996-1004, :1022-1070
SW_DEV LmE lm_mix(const LmE& a, const LmE& b, float ratio, bool directional)
{
	const float omr = 1.0f - ratio;
	LmE r;
	r.amb[0] = mul_add(a.amb[0], ratio, b.amb[0] * omr);
	r.amb[1] = mul_add(a.amb[1], ratio, b.amb[1] * omr);
	r.amb[2] = mul_add(a.amb[2], ratio, b.amb[2] * omr);
	if(directional)
	{
		r.dir = v3(mul_add(a.dir.x, ratio, b.dir.x * omr), mul_add(a.dir.y, ratio, b.dir.y * omr),
			mul_add(a.dir.z, ratio, b.dir.z * omr));
		r.dev = mul_add(a.dev, ratio, b.dev * omr);
		r.col[0] = mul_add(a.col[0], ratio, b.col[0] * omr);
		r.col[1] = mul_add(a.col[1], ratio, b.col[1] * omr);
		r.col[2] = mul_add(a.col[2], ratio, b.col[2] * omr);
	}
	else
	{
		r.dir = v3(0.0f, 0.0f, 0.0f);
		r.dev = 0.0f;
		r.col[0] = r.col[1] = r.col[2] = 0.0f;
	}
	return r;
}

// A lightmap line element in LDS: 12 floats for a directional element (ambient, direction,
// deviation, colour, two unused), 4 for a simple one (ambient, one unused), so every element
// starts on 16 bytes.
SW_DEV uint32_t line_stride(bool directional) { return directional ? 12u : 4u; }

SW_DEV void line_store(float* p, const LmE& e, bool directional)
{
	if(directional)
	{
		*reinterpret_cast<float4*>(p) = make_float4(e.amb[0], e.amb[1], e.amb[2], e.dir.x);
		*reinterpret_cast<float4*>(p + 4) = make_float4(e.dir.y, e.dir.z, e.dev, e.col[0]);
		*reinterpret_cast<float2*>(p + 8) = make_float2(e.col[1], e.col[2]);
	}
	else
		*reinterpret_cast<float4*>(p) = make_float4(e.amb[0], e.amb[1], e.amb[2], 0.0f);
}

SW_DEV LmE line_load(const float* p, bool directional)
{
	LmE e;
	if(directional)
	{
		const float4 a = *reinterpret_cast<const float4*>(p);
		const float4 b = *reinterpret_cast<const float4*>(p + 4);
		const float2 c = *reinterpret_cast<const float2*>(p + 8);
		e.amb[0] = a.x;
		e.amb[1] = a.y;
		e.amb[2] = a.z;
		e.dir = v3(a.w, b.x, b.y);
		e.dev = b.z;
		e.col[0] = b.w;
		e.col[1] = c.x;
		e.col[2] = c.y;
	}
	else
	{
		const float4 a = *reinterpret_cast<const float4*>(p);
		e.amb[0] = a.x;
		e.amb[1] = a.y;
		e.amb[2] = a.z;
		e.dir = v3(0.0f, 0.0f, 0.0f);
		e.dev = 0.0f;
		e.col[0] = e.col[1] = e.col[2] = 0.0f;
	}
	return e;
}

// surfaces.rs:1088-1150
SW_DEV float specular_intensity(float c, float inv_roughness)
{
	const float x = (c - 1.0f) * inv_roughness;
	const float SQRT_7 = 2.645751311f;
	const float A = 2.0f * 3.14159265358979323846f / SQRT_7;
	const float B = 3.14159265358979323846f * SQRT_7 / 2.0f;
	return inv_fast(mul_add(x * x, B, A)) * inv_roughness;
}
SW_DEV float fresnel_factor_base(float c)
{
	const float a = fmax_r(1.0f - c, 0.0f);
	const float a2 = a * a;
	return a2 * a2 * a;
}
SW_DEV float specular_k(uint32_t spec, float f_base, float roughness)
{
	if(spec == 1)
	{
		const float f = mul_add(f_base, 1.0f - 0.04f, 0.04f);
		return mul_add(f, 1.0f - roughness, (0.04f * 3.0f) * roughness);
	}
	return mul_add(f_base, 1.0f - roughness, 0.5f * roughness);
}

// Per-surface constants of the texel loop: what surfaces.rs:433-470 and :903-915 compute once per
// surface (basis normalisation, start position, texture wrap, lightmap steps, emissive light
// vector), computed once per surface and block by the surface's lane of the block's first wave and
// kept in LDS for every texel of the block.
struct SurfRec
{
	uint64_t start, end; // arena texel range [start, end)
	uint64_t tex_base; // m.texels index of the diffuse level's texel (0, 0)
	uint64_t etex_base; // m.lite index of the emissive level's texel (0, 0)
	double inv_w, inv_tw, inv_th, inv_ew, inv_eh; // double reciprocals of the divisors below
	uint32_t w; // size[0]
	uint32_t flags; // REC_*
	uint32_t tw, th, ew, eh; // diffuse and emissive level sizes
	uint32_t tw_mask, th_mask, ew_mask, eh_mask; // size - 1 for a power of two, else NO_MASK
	int32_t tcu, tcv, etcu, etcv; // tc_start reduced modulo the level sizes (as given with *_SLOW_WRAP)
	uint32_t first_light;
	uint32_t lm_offset, lm_stride, lm_sl, lm_su, lm_sv;
	int32_t elight[3]; // cvtps2dq(emissive_light * 256), lanes B, G, R
	float inv_scale, k_shift;
	V3 start_pos, u_vec, v_vec, u_n, v_n, normal, cam;
	// The surface's lightmap lines in the block's LDS (blocks that stage them): line (row, column)
	// = the vertical mix of lightmap rows lv and lv + 1 for surface row line_dv0 + row, at lightmap
	// column line_lu0 + column; line_off (in floats) is NO_LINES when the surface's texels fetch
	// their four elements from global memory instead.
	uint32_t line_off, line_e0, line_ne, line_dv0, line_lu0, line_ncols;
};

enum : uint32_t
{
	REC_DIFFUSE = 1u,
	REC_EMISSIVE = 2u,
	REC_NORMAL_MAP = 4u,
	REC_DIRECTIONAL = 8u,
	REC_LIGHTMAP = 16u,
	REC_SKIP = 32u, // invalid descriptor, or neither layer: the texels are left untouched
	REC_SLOW_WRAP = 64u, // sizes or coordinates outside the fast wrap's range
	REC_EMISSIVE_SLOW_WRAP = 128u,
	REC_SPEC_SHIFT = 8, // 2 bits: 0 none, 1 dielectric, 2 metal
	REC_BLEND_SHIFT = 12, // 4 bits: the emissive layer's blend
	REC_LIGHTS_SHIFT = 16, // clamped light count
};
#define SURF_CHUNK 32 // surface records per pass of a block
#define LINE_FLOATS 4096 // LDS floats of a block's staged lightmap lines (16 KiB)
#define NO_LINES 0xFFFFFFFFu
#define SURF_LINES_MIN_TEXELS 1024 // blocks of at least this many texels stage lightmap lines

using swdiv::div_rcp;
using swdiv::mod_rcp;
using swdiv::NO_MASK;
using swdiv::pow2_mask;

// Wrap set-up of one texture level: with tc_start[1] + dst_v free of i32 overflow and sizes below
// 2^30, rem_euclid(tc_start + d, size) == (rem_euclid(tc_start, size) + d) mod size, which the
// texel evaluates with mod_rcp; otherwise the texel keeps the reference's expressions.
SW_DEV bool wrap_setup(const sw_surface_desc& d, uint32_t tw, uint32_t th, int32_t& tcu, int32_t& tcv, uint32_t& tw_mask,
	uint32_t& th_mask, double& inv_tw, double& inv_th)
{
	const bool fast = tw >= 1u && th >= 1u && tw <= (1u << 30) && th <= (1u << 30) &&
		int64_t(d.tc_start[1]) + int64_t(d.size[1]) <= int64_t(INT32_MAX) + 1;
	if(fast)
	{
		tcu = rem_euclid(d.tc_start[0], int32_t(tw));
		tcv = rem_euclid(d.tc_start[1], int32_t(th));
	}
	else
	{
		tcu = d.tc_start[0];
		tcv = d.tc_start[1];
	}
	tw_mask = pow2_mask(tw);
	th_mask = pow2_mask(th);
	inv_tw = 1.0 / double(tw);
	inv_th = 1.0 / double(th);
	return fast;
}

// Fills the record of a surface (ok: the descriptor was validated).
SW_DEV void make_surf_rec(SurfRec& r, const sw_surface_desc& d, uint64_t start, uint64_t end, const sw_view& view,
	const MapPtrs& m, bool ok)
{
	r.start = start;
	r.end = end;
	r.line_off = NO_LINES;
	r.line_e0 = r.line_ne = 0;
	r.w = d.size[0];
	r.inv_w = 1.0 / double(d.size[0] ? d.size[0] : 1u);
	const bool diffuse = (d.flags & SW_SURF_DIFFUSE_LAYER) != 0;
	const bool emissive = (d.flags & SW_SURF_EMISSIVE) && d.emissive_level < m.num_levels;
	uint32_t fl = 0;
	if(!ok || (!diffuse && !emissive))
	{
		r.flags = REC_SKIP;
		return;
	}
	if(diffuse)
	{
		const sw_texture_level tex = m.levels[d.texture_level];
		fl |= REC_DIFFUSE;
		if(tex.flags & SW_TEX_HAS_NORMAL_MAP)
			fl |= REC_NORMAL_MAP;
		const uint32_t spec = (tex.flags & SW_TEX_HAS_NON_ONE_ROUGHNESS) ? ((tex.flags & SW_TEX_IS_METAL) ? 2u : 1u) : 0u;
		fl |= spec << REC_SPEC_SHIFT;
		if(d.flags & SW_SURF_DIRECTIONAL)
			fl |= REC_DIRECTIONAL;
		if(d.lightmap_offset != 0)
			fl |= REC_LIGHTMAP;
		const uint32_t nl = d.num_lights < SW_MAX_POLYGON_LIGHTS ? d.num_lights : SW_MAX_POLYGON_LIGHTS;
		fl |= nl << REC_LIGHTS_SHIFT;
		r.tex_base = tex.offset;
		r.tw = tex.width;
		r.th = tex.height;
		if(!wrap_setup(d, tex.width, tex.height, r.tcu, r.tcv, r.tw_mask, r.th_mask, r.inv_tw, r.inv_th))
			fl |= REC_SLOW_WRAP;

		const V3 u_vec = v3(d.basis_u[0], d.basis_u[1], d.basis_u[2]);
		const V3 v_vec = v3(d.basis_v[0], d.basis_v[1], d.basis_v[2]);
		r.u_vec = u_vec;
		r.v_vec = v_vec;
		r.normal = v3(d.basis_normal[0], d.basis_normal[1], d.basis_normal[2]);
		// :456-457
		r.start_pos = add(add(v3(d.basis_start[0], d.basis_start[1], d.basis_start[2]), scale(u_vec, float(d.tc_start[0]) + 0.5f)),
			scale(v_vec, float(d.tc_start[1]) + 0.5f));
		r.u_n = scale(u_vec, inv_sqrt_fast(fmax_r(len2(u_vec), MIN_POSITIVE_VALUE)));
		r.v_n = scale(v_vec, inv_sqrt_fast(fmax_r(len2(v_vec), MIN_POSITIVE_VALUE)));
		r.cam = v3(view.cam_pos[0], view.cam_pos[1], view.cam_pos[2]);

		const uint32_t sl = d.lightmap_scale_log2;
		r.lm_sl = sl;
		r.inv_scale = 1.0f / float(1u << sl);
		r.k_shift = 0.5f * r.inv_scale;
		r.lm_offset = d.lightmap_offset;
		r.lm_stride = d.lightmap_size[0];
		r.lm_su = d.lightmap_tc_shift[0];
		r.lm_sv = d.lightmap_tc_shift[1];
		r.first_light = d.first_light;
	}
	if(emissive)
	{
		const sw_texture_level tex = m.levels[d.emissive_level];
		// blends past the known ones keep the destination (the switch's default), as 15 does
		const uint32_t blend = d.emissive_blend < 15u ? d.emissive_blend : 15u;
		fl |= REC_EMISSIVE | (blend << REC_BLEND_SHIFT);
		r.etex_base = tex.offset;
		r.ew = tex.width;
		r.eh = tex.height;
		if(!wrap_setup(d, tex.width, tex.height, r.etcu, r.etcv, r.ew_mask, r.eh_mask, r.inv_ew, r.inv_eh))
			fl |= REC_EMISSIVE_SLOW_WRAP;
		// light_vec = cvtps2dq(light * 256): lanes [B, G, R]
		r.elight[0] = f32_to_i32_rne_x86(d.emissive_light[2] * 256.0f);
		r.elight[1] = f32_to_i32_rne_x86(d.emissive_light[1] * 256.0f);
		r.elight[2] = f32_to_i32_rne_x86(d.emissive_light[0] * 256.0f);
	}
	r.flags = fl;
}

// Texel coordinates of (dst_u, dst_v) in a wrapped level (:507-517, :843-847): rem_euclid of
// tc_start + dst per axis.
SW_DEV void wrap_texel(bool slow, int32_t tcu, int32_t tcv, uint32_t tw, uint32_t th, uint32_t tw_mask, uint32_t th_mask,
	double inv_tw, double inv_th, uint32_t dst_u, uint32_t dst_v, uint32_t& src_u, uint32_t& src_v)
{
	if(slow)
	{
		src_v = uint32_t(rem_euclid(tcv + int32_t(dst_v), int32_t(th)));
		src_u = uint32_t(rem_euclid(rem_euclid(tcu, int32_t(tw)) + int32_t(dst_u % tw), int32_t(tw)));
	}
	else
	{
		src_v = mod_rcp(uint32_t(tcv) + dst_v, th, th_mask, inv_th);
		src_u = mod_rcp(uint32_t(tcu) + dst_u, tw, tw_mask, inv_tw);
	}
}

// cube_shadow_map_fetch / cube_shadow_map_side_fetch (light.rs:195-257) and
// projector_shadow_map_fetch (:259-294): true when the point at -vec from the light is lit (its
// 1 / z is at least the map's). The map lies inside the shadow texels (checked on the host).
SW_DEV bool shadow_lit(const sw_light_shadow& S, const float* __restrict__ texels, V3 vec)
{
	if(S.kind == SW_SHADOW_CUBE)
	{
		const V3 a = v3(__builtin_fabsf(vec.x), __builtin_fabsf(vec.y), __builtin_fabsf(vec.z));
		V3 v;
		uint32_t side;
		if(a.x >= a.y && a.x >= a.z)
		{
			v = vec.x >= 0.0f ? v3(-vec.y, vec.z, a.x) : v3(vec.y, vec.z, a.x);
			side = vec.x >= 0.0f ? 1u : 0u;
		}
		else if(a.y >= a.x && a.y >= a.z)
		{
			v = vec.y >= 0.0f ? v3(vec.x, vec.z, a.y) : v3(-vec.x, vec.z, a.y);
			side = vec.y >= 0.0f ? 3u : 2u;
		}
		else
		{
			v = vec.z >= 0.0f ? v3(-vec.x, vec.y, a.z) : v3(-vec.x, -vec.y, a.z);
			side = vec.z >= 0.0f ? 5u : 4u;
		}
		const float one_minus_eps = 1.0f - 1.0f / 65536.0f;
		const float size_f = float(S.size);
		const float depth = inv_fast(fmax_r(v.z, MIN_POSITIVE_VALUE));
		const float half_depth = 0.5f * depth;
		const float u_f = fmin_r(fmax_r(mul_add(v.x, half_depth, 0.5f), 0.0f), one_minus_eps) * size_f;
		const float v_f = fmin_r(fmax_r(mul_add(v.y, half_depth, 0.5f), 0.0f), one_minus_eps) * size_f;
		const uint32_t u = uint32_t(u_f), w = uint32_t(v_f);
		const uint64_t side_texels = uint64_t(S.size) * S.size;
		return depth >= texels[S.offset + side * side_texels + uint64_t(w) * S.size + u];
	}
	if(S.kind == SW_SHADOW_PROJECTOR)
	{
		const V3 bx = v3(S.basis_x[0], S.basis_x[1], S.basis_x[2]), by = v3(S.basis_y[0], S.basis_y[1], S.basis_y[2]),
				 bz = v3(S.basis_z[0], S.basis_z[1], S.basis_z[2]);
		const float z = dot(bz, vec);
		if(z <= 0.0f)
			return false;
		const float depth = inv_fast(z);
		const float half_depth = 0.5f * depth;
		const float size_f = float(S.size);
		const float u_f = mul_add(dot(bx, vec), half_depth, 0.5f) * size_f;
		const float v_f = mul_add(dot(by, vec), half_depth, 0.5f) * size_f;
		if(!(u_f >= 0.0f && v_f >= 0.0f && u_f < size_f && v_f < size_f))
			return false; // outside (a NaN coordinate too)
		return depth >= texels[S.offset + uint64_t(uint32_t(v_f)) * S.size + uint32_t(u_f)];
	}
	return true;
}

// One texel of build_surface_impl_6_static_params (surfaces.rs:433-852), specialisations as flags
// that are uniform per surface. SH: lights may have shadow maps (a separate instantiation, so the
// kernel without them keeps its registers).
template <class CT, bool SH, bool LINES>
SW_DEV typename CT::T surface_texel(const MapPtrs& m, const SurfRec& R, uint32_t fl, const sw_light* __restrict__ lights,
	const uint32_t* __restrict__ light_indices, uint32_t num_light_indices, uint32_t num_lights_total, uint32_t dst_u,
	uint32_t dst_v, const SurfaceArgs& a, const float* lines)
{
	const bool use_normal_map = (fl & REC_NORMAL_MAP) != 0;
	const uint32_t spec = (fl >> REC_SPEC_SHIFT) & 3u;
	const bool directional = (fl & REC_DIRECTIONAL) != 0;
	const bool has_lightmap = (fl & REC_LIGHTMAP) != 0;
	const uint32_t num_lights = fl >> REC_LIGHTS_SHIFT;

	uint32_t src_u, src_v;
	wrap_texel((fl & REC_SLOW_WRAP) != 0, R.tcu, R.tcv, R.tw, R.th, R.tw_mask, R.th_mask, R.inv_tw, R.inv_th, dst_u, dst_v,
		src_u, src_v);
	sw_texel texel;
	{
		// the 8-byte texel as one 64-bit load (diffuse in the low half)
		const uint64_t t = elem32(reinterpret_cast<const uint64_t*>(m.texels), uint32_t(R.tex_base) + src_v * R.tw + src_u);
		texel.diffuse = uint32_t(t);
		texel.packed_normal_roughness = int32_t(uint32_t(t >> 32));
	}

	const V3 u_vec = R.u_vec, v_vec = R.v_vec;
	const V3 start_pos_v = scalar_mul_add(v_vec, float(dst_v), R.start_pos);
	const V3 pos = scalar_mul_add(u_vec, float(dst_u), start_pos_v);

	V3 tn;
	float tr;
	if(use_normal_map)
	{
		const float k = 1.0f / 127.0f;
		const int32_t p = texel.packed_normal_roughness;
		tn = v3(float(int32_t(int8_t(p & 0xFF))) * k, float(int32_t(int8_t((p >> 8) & 0xFF))) * k,
			float(int32_t(int8_t((p >> 16) & 0xFF))) * k);
		tr = float(p >> 24) * k;
	}
	else
	{
		tn = v3(0.0f, 0.0f, 1.0f);
		tr = spec == 0 ? 0.0f : float(texel.packed_normal_roughness >> 24) * (1.0f / 127.0f);
	}

	CV albedo_mod = cv_zero();
	CV direct = cv_zero();
	if(has_lightmap)
	{
		const uint32_t sl = R.lm_sl;
		const float inv_scale = R.inv_scale;
		const float k_shift = R.k_shift;
		const uint32_t mask = (1u << sl) - 1u;
		const uint32_t base_u = dst_u + R.lm_su;
		const uint32_t lu = base_u >> sl;
		const float ku = mul_add(float(base_u & mask), inv_scale, k_shift);
		LmE line0, line1;
		if(LINES && R.line_off != NO_LINES)
		{
			// the row's lightmap line (:473-502), mixed once per block and column
			const uint32_t stride = line_stride(directional);
			const float* p = lines + R.line_off + ((dst_v - R.line_dv0) * R.line_ncols + (lu - R.line_lu0)) * stride;
			line0 = line_load(p, directional);
			line1 = line_load(p + stride, directional);
		}
		else
		{
			// vertical mix of the two lightmap rows (:477-502) for the two columns this texel needs
			const uint32_t base_v = dst_v + R.lm_sv;
			const uint32_t lv = base_v >> sl;
			const float kv = mul_add(float(base_v & mask), inv_scale, k_shift);
			const uint32_t row0 = R.lm_offset + lv * R.lm_stride;
			const uint32_t row1 = row0 + R.lm_stride;
			const LmE l00 = lm_load(m, directional, row0 + lu);
			const LmE l01 = lm_load(m, directional, row0 + lu + 1);
			const LmE l10 = lm_load(m, directional, row1 + lu);
			const LmE l11 = lm_load(m, directional, row1 + lu + 1);
			line0 = lm_mix(l10, l00, kv, directional);
			line1 = lm_mix(l11, l01, kv, directional);
		}
		const LmE lm = lm_mix(line1, line0, ku, directional); // (:544-549)

		if(spec == 0)
		{
			albedo_mod = cv_rgb1(lm.amb[0], lm.amb[1], lm.amb[2]);
			if(directional)
			{
				const float d = use_normal_map ? fmax_r(dot(lm.dir, tn), 0.0f) : lm.dir.z;
				albedo_mod = cv_mul_scalar_add(cv_rgb(lm.col[0], lm.col[1], lm.col[2]), d, albedo_mod);
			}
		}
		else
		{
			const V3 normal = use_normal_map ? tn : v3(0.0f, 0.0f, 1.0f);
			const V3 to_cam = sub(R.cam, pos);
			const V3 to_cam_ts = v3(dot(to_cam, R.u_n), dot(to_cam, R.v_n), dot(to_cam, R.normal));
			const float n_dot = dot(to_cam_ts, normal);
			const V3 refl = sub(scale(normal, 2.0f * n_dot), to_cam_ts);
			const float refl_len2 = fmax_r(len2(refl), MIN_POSITIVE_VALUE);
			const float cos_a = fmax_r(n_dot * inv_sqrt_fast(refl_len2), 0.0f);
			const float sk = specular_k(spec, fresnel_factor_base(cos_a), tr);
			const float omsk = 1.0f - sk;
			const CV cc = cv_rgb(lm.amb[0], lm.amb[1], lm.amb[2]);
			albedo_mod = cv_scalar_mul(cc, omsk);
			direct = cv_scalar_mul(cc, sk);
			albedo_mod.a = 1.0f;
			if(directional)
			{
				const float dl2 = len2(lm.dir);
				const float dl = dl2 * inv_sqrt_fast(dl2);
				const float rc = dot(refl, lm.dir) * inv_sqrt_fast(refl_len2 * dl2);
				const float inv_r = fmax_r(inv_fast(tr + lm.dev), 0.75f);
				const float si = specular_intensity(rc, inv_r);
				const CV dc = cv_rgb(lm.col[0], lm.col[1], lm.col[2]);
				if(spec == 1)
				{
					const float di = fmax_r(dot(lm.dir, tn), 0.0f);
					albedo_mod = cv_mul_scalar_add(dc, di * omsk, albedo_mod);
					direct = cv_mul_scalar_add(dc, si * sk * dl, direct);
				}
				else
				{
					const float sf = si * dl;
					albedo_mod = cv_mul_scalar_add(dc, omsk * sf, albedo_mod);
					direct = cv_mul_scalar_add(dc, sk * sf, direct);
				}
			}
		}
	}

	if(num_lights > 0)
	{
		// :683-831
		const V3 u_n = R.u_n, v_n = R.v_n, basis_normal = R.normal;
		V3 normal;
		if(use_normal_map)
			normal = v3(mul_add(tn.x, u_n.x, mul_add(tn.y, v_n.x, tn.z * basis_normal.x)),
				mul_add(tn.x, u_n.y, mul_add(tn.y, v_n.y, tn.z * basis_normal.y)),
				mul_add(tn.x, u_n.z, mul_add(tn.y, v_n.z, tn.z * basis_normal.z)));
		else
			normal = basis_normal;

		V3 refl = v3(0.0f, 0.0f, 0.0f);
		float refl_len2 = MIN_POSITIVE_VALUE, sk = 0.0f, inv_r = 0.0f;
		if(spec != 0)
		{
			const V3 to_cam = sub(R.cam, pos);
			const float n_dot = dot(to_cam, normal);
			refl = sub(scale(normal, 2.0f * n_dot), to_cam);
			refl_len2 = fmax_r(len2(refl), MIN_POSITIVE_VALUE);
			const float cos_a = fmax_r(n_dot * inv_sqrt_fast(refl_len2), 0.0f);
			sk = specular_k(spec, fresnel_factor_base(cos_a), tr);
			inv_r = inv_fast(tr);
		}
		const uint32_t first_light = R.first_light;
		for(uint32_t li = 0; li < num_lights; ++li)
		{
			const uint32_t idx_pos = first_light + li;
			if(idx_pos >= num_light_indices)
				break;
			const uint32_t lidx = light_indices[idx_pos];
			if(lidx >= num_lights_total)
				continue;
			const sw_light L = lights[lidx];
			const V3 to_light = sub(v3(L.position[0], L.position[1], L.position[2]), pos);
			const float tl2 = fmax_r(len2(to_light), MIN_POSITIVE_VALUE);
			const float dist_f = fmax_r(inv_fast(tl2) - L.inv_square_radius, 0.0f);
			// shadow factor 1 (ShadowMap::None or lit) or 0; dist_f is finite and >= 0, so 0 * dist_f is +0
			float sdf = 1.0f * dist_f;
			if(SH && lidx < a.num_shadows && !shadow_lit(a.shadows[lidx], a.shadow_texels, to_light))
				sdf = 0.0f;
			const float di = spec == 2 ? 0.0f : fmax_r(dot(normal, to_light) * inv_sqrt_fast(tl2), 0.0f);
			float si = 0.0f;
			if(spec != 0)
				si = specular_intensity(dot(refl, to_light) * inv_sqrt_fast(refl_len2 * tl2), inv_r);
			const CV lc = cv_rgb(L.color[0], L.color[1], L.color[2]);
			if(spec == 0)
				albedo_mod = cv_mul_scalar_add(lc, di * sdf, albedo_mod);
			else if(spec == 1)
			{
				albedo_mod = cv_mul_scalar_add(lc, di * (1.0f - sk) * sdf, albedo_mod);
				direct = cv_mul_scalar_add(lc, si * sk * sdf, direct);
			}
			else
			{
				const float x = si * sdf;
				albedo_mod = cv_mul_scalar_add(lc, (1.0f - sk) * x, albedo_mod);
				direct = cv_mul_scalar_add(lc, sk * x, direct);
			}
		}
	}

	// :833-842
	CV result = cv_mul(cv_from_color32(texel.diffuse), albedo_mod);
	if(spec != 0)
		result = cv_mul_scalar_add(direct, 255.0f, result);
	return CT::from_cv(result);
}

template <class CT>
SW_DEV typename CT::T emissive_blend(uint32_t t, int32_t lb, int32_t lg, int32_t lr, uint32_t blend, typename CT::T dst);

// surfaces.rs:903-971, one texel
template <class CT>
SW_DEV typename CT::T emissive_texel(const MapPtrs& m, const SurfRec& R, uint32_t fl, typename CT::T dst, uint32_t dst_u,
	uint32_t dst_v)
{
	uint32_t src_u, src_v;
	wrap_texel((fl & REC_EMISSIVE_SLOW_WRAP) != 0, R.etcu, R.etcv, R.ew, R.eh, R.ew_mask, R.eh_mask, R.inv_ew, R.inv_eh,
		dst_u, dst_v, src_u, src_v);
	const uint32_t t = elem32(m.lite, uint32_t(R.etex_base) + src_v * R.ew + src_u);
	return emissive_blend<CT>(t, R.elight[0], R.elight[1], R.elight[2], (fl >> REC_BLEND_SHIFT) & 15u, dst);
}

// surfaces.rs:920-968: the lite texel t modulated by the integer light (lanes B, G, R; alpha by 0)
// and blended into dst
template <class CT>
SW_DEV typename CT::T emissive_blend(uint32_t t, int32_t lb, int32_t lg, int32_t lr, uint32_t blend, typename CT::T dst)
{
	const CI tc = Color32T::to_ci(t);
	const CI mod = CI{wmul(tc.b, lb) >> 8, wmul(tc.g, lg) >> 8, wmul(tc.r, lr) >> 8, wmul(tc.a, 0) >> 8};
	switch(blend)
	{
	case SW_BLEND_NONE: return CT::from_ci(mod);
	case SW_BLEND_AVERAGE:
	{
		const CI d = CT::to_ci(dst);
		return CT::from_ci(CI{wadd(mod.b, d.b) >> 1, wadd(mod.g, d.g) >> 1, wadd(mod.r, d.r) >> 1, wadd(mod.a, d.a) >> 1});
	}
	case SW_BLEND_ADDITIVE:
	{
		const CI d = CT::to_ci(dst);
		return CT::from_ci(CI{wadd(mod.b, d.b), wadd(mod.g, d.g), wadd(mod.r, d.r), wadd(mod.a, d.a)});
	}
	case SW_BLEND_ALPHA_TEST: return Color32T::test_alpha(t) ? CT::from_ci(mod) : dst;
	case SW_BLEND_ALPHA_BLEND:
	{
		const int32_t alpha = Color32T::get_alpha(t);
		const CI d = CT::to_ci(dst);
		const int32_t ia = 255 - alpha;
		return CT::from_ci(CI{wadd(wmul(mod.b, alpha), wmul(d.b, ia)) >> 8, wadd(wmul(mod.g, alpha), wmul(d.g, ia)) >> 8,
			wadd(wmul(mod.r, alpha), wmul(d.r, ia)) >> 8, wadd(wmul(mod.a, alpha), wmul(d.a, ia)) >> 8});
	}
	default: return dst;
	}
}

// One block of surface texels: surf_block_texels() consecutive arena texels, up to
// SURF_MAX_TEXELS_PER_THREAD per thread (strided by the block size, so each pass is coalesced). The block's first wave builds
// the records of up to SURF_CHUNK consecutive surfaces overlapping the block (one lane each), then
// every thread finds its texels' surfaces among them by a forward scan; a block overlapping more
// surfaces takes further passes. LINES: before the texels, the block mixes each lightmapped
// surface's lightmap lines for its rows (as the reference does per row, surfaces.rs:473-502) into
// LDS, up to LINE_FLOATS; the texels then read two line elements from LDS instead of four
// lightmap elements from global memory (same operations in the same order, so the same bits).
template <class CT, bool SH, bool LINES>
SW_DEV void build_surfaces_block(uint32_t block, const MapPtrs& m, const SurfaceArgs& a, const ViewsArg& views)
{
	__shared__ SurfRec recs[SURF_CHUNK];
	__shared__ uint32_t sh_base, sh_count, sh_lines;
	__shared__ __align__(16) float lines[LINES ? LINE_FLOATS : 4];
	const sw_surface_desc* __restrict__ surfaces = a.surfaces;
	const uint32_t num_surfaces = a.num_surfaces;
	const SurfExt* __restrict__ surf_ext = a.surf_ext;
	typename CT::T* __restrict__ arena = static_cast<typename CT::T*>(a.arena);
	const uint32_t bt = surf_block_texels(a.arena_texels);
	const uint64_t b0 = uint64_t(block) * bt;
	if(b0 >= a.arena_texels || num_surfaces == 0)
		return; // uniform over the block
	const uint64_t b1 = std::min<uint64_t>(b0 + bt, a.arena_texels);
	const uint32_t tid = threadIdx.x;
	const bool search = a.blk_surf == nullptr;
	const uint64_t search_base = search ? views[0].v.surface_arena_offset : 0;
	if(tid < 64)
	{
		uint32_t first;
		if(!search)
			first = a.blk_surf[block]; // the surface holding b0, or the next one after a gap
		else
		{
			// One view, no table: a 64-ary search by the wave (one probe per lane and level) finds
			// the last surface starting at or before b0. The descriptors were not validated ahead,
			// so each record's lane validates its surface below.
			const uint32_t lane = tid;
			uint32_t lo = 0, hi = num_surfaces;
			while(hi - lo > 1)
			{
				const uint32_t step = (hi - lo + 63) / 64;
				const uint32_t probe = lo + lane * step;
				const bool le = probe < hi && search_base + surfaces[probe].pixels_offset <= b0;
				const uint64_t mask = __ballot(le);
				const uint32_t last = mask ? 63u - uint32_t(__builtin_clzll(mask)) : 0u;
				lo = lo + last * step;
				hi = std::min(lo + step, hi);
			}
			first = lo;
		}
		if(tid == 0)
			sh_base = first;
	}
	__syncthreads();
	uint32_t base = sh_base;
	for(;;)
	{
		if(tid < 64)
		{
			const uint32_t s = base + tid;
			bool overlaps = false;
			if(tid < SURF_CHUNK && s < num_surfaces)
			{
				const sw_surface_desc& d = surfaces[s];
				const uint64_t size = uint64_t(d.size[0]) * d.size[1];
				uint64_t start, end;
				uint32_t view = 0;
				bool ok = true;
				if(!search)
				{
					const SurfExt e = surf_ext[s];
					end = e.end;
					start = end - size;
					view = e.view;
				}
				else
				{
					start = search_base + d.pixels_offset;
					end = start + size;
				}
				overlaps = start < b1;
				if(overlaps)
				{
					if(search)
					{
						ok = surface_desc_ok(d, end, views[0].arena_end, m);
						if(s > 0)
						{
							const sw_surface_desc& p = surfaces[s - 1];
							ok = ok && search_base + p.pixels_offset + uint64_t(p.size[0]) * p.size[1] <= start;
						}
						if(!ok)
							raise_error(a.counters, a.host_err, ERROR_BAD_SURFACE);
					}
					make_surf_rec(recs[tid], d, start, end, views[view].v, m, ok);
				}
			}
			// surfaces follow arena order: the overlapping ones are a prefix of the chunk
			const uint64_t mask = __ballot(overlaps);
			if(tid == 0)
				sh_count = uint32_t(__builtin_ctzll(~mask));
		}
		__syncthreads();
		const uint32_t count = sh_count;
		if(LINES)
		{
			if(tid < 64)
			{
				// the lines each surface's rows in [b0, b1) need: every column a full row reads
				uint32_t ne = 0, nf = 0, dv0 = 0, lu0 = 0, ncols = 0;
				if(tid < count)
				{
					const SurfRec& R = recs[tid];
					const uint64_t lo = std::max<uint64_t>(R.start, b0), hi = std::min<uint64_t>(R.end, b1);
					if((R.flags & (REC_SKIP | REC_DIFFUSE | REC_LIGHTMAP)) == (REC_DIFFUSE | REC_LIGHTMAP) && lo < hi)
					{
						uint32_t rem;
						dv0 = div_rcp(uint32_t(lo - R.start), R.w, R.inv_w, rem);
						const uint32_t dv1 = div_rcp(uint32_t(hi - 1 - R.start), R.w, R.inv_w, rem);
						lu0 = R.lm_su >> R.lm_sl;
						ncols = ((R.w - 1u + R.lm_su) >> R.lm_sl) - lu0 + 2u;
						// rows <= bt / w + 2 and columns <= w + 2: at most bt + 2w + 2bt/w + 4 < 2^18
						// elements per record, so the sums below stay far below 2^32
						ne = (dv1 - dv0 + 1u) * ncols;
						nf = ne * line_stride((R.flags & REC_DIRECTIONAL) != 0);
					}
				}
				uint32_t sf = nf, se = ne; // inclusive prefix sums over the records
#pragma unroll
				for(uint32_t dlt = 1; dlt < 64; dlt <<= 1)
				{
					const uint32_t tf = __shfl_up(sf, dlt), te = __shfl_up(se, dlt);
					if(tid >= dlt)
					{
						sf += tf;
						se += te;
					}
				}
				// the staged records are a prefix (skipping those without lines): sf never decreases
				const bool fits = ne > 0 && sf <= LINE_FLOATS;
				if(tid < count)
				{
					SurfRec& R = recs[tid];
					R.line_off = fits ? sf - nf : NO_LINES;
					R.line_e0 = se - ne;
					R.line_ne = fits ? ne : 0u;
					R.line_dv0 = dv0;
					R.line_lu0 = lu0;
					R.line_ncols = ncols;
				}
				const uint64_t fm = __ballot(fits);
				const uint32_t total = __shfl(se, fm ? 63 - __builtin_clzll(fm) : 0);
				if(tid == 0)
					sh_lines = fm ? total : 0u;
			}
			__syncthreads();
			const uint32_t total = sh_lines;
			uint32_t lr = 0;
#pragma unroll 1
			for(uint32_t j = tid; j < total; j += SURF_BLOCK)
			{
				// j < total lies in a staged record; those without lines have line_ne 0
				while(j >= recs[lr].line_e0 + recs[lr].line_ne)
					++lr;
				const SurfRec& R = recs[lr];
				const uint32_t q = j - R.line_e0;
				const uint32_t row = q / R.line_ncols;
				const uint32_t col = q - row * R.line_ncols;
				const bool directional = (R.flags & REC_DIRECTIONAL) != 0;
				const uint32_t sl = R.lm_sl;
				// :477-502 for surface row line_dv0 + row and lightmap column line_lu0 + col
				const uint32_t base_v = R.line_dv0 + row + R.lm_sv;
				const uint32_t lv = base_v >> sl;
				const float kv = mul_add(float(base_v & ((1u << sl) - 1u)), R.inv_scale, R.k_shift);
				const uint32_t row0 = R.lm_offset + lv * R.lm_stride;
				const uint32_t lu = R.line_lu0 + col;
				const LmE l0 = lm_load(m, directional, row0 + lu);
				const LmE l1 = lm_load(m, directional, row0 + R.lm_stride + lu);
				line_store(lines + R.line_off + q * line_stride(directional), lm_mix(l1, l0, kv, directional), directional);
			}
			__syncthreads();
		}
		uint32_t r = 0;
#pragma unroll 1
		for(uint32_t k = 0; k < bt / SURF_BLOCK; ++k)
		{
			const uint64_t idx = b0 + uint64_t(k) * SURF_BLOCK + tid;
			if(idx >= b1)
				break;
			while(r < count && recs[r].end <= idx)
				++r;
			if(r == count)
				break; // this and later texels lie past the pass's surfaces
			const SurfRec& R = recs[r];
			const uint32_t fl = R.flags;
			if(idx < R.start || (fl & REC_SKIP))
				continue; // gap between surfaces, or a surface left untouched
			uint32_t du;
			const uint32_t dv = div_rcp(uint32_t(idx - R.start), R.w, R.inv_w, du);
			typename CT::T out = 0;
			if(fl & REC_DIFFUSE)
				out = surface_texel<CT, SH, LINES>(
					m, R, fl, a.lights, a.light_indices, a.num_light_indices, a.num_lights, du, dv, a, lines);
			else if(((fl >> REC_BLEND_SHIFT) & 15u) != SW_BLEND_NONE)
				out = arena[idx]; // an emissive layer without a diffuse layer blends over what the arena holds
			if(fl & REC_EMISSIVE)
				out = emissive_texel<CT>(m, R, fl, out, du, dv);
			arena[idx] = out;
		}
		if(count < SURF_CHUNK)
			break;
		__syncthreads(); // the next pass rewrites the records
		base += SURF_CHUNK;
	}
}

// Held to 80 VGPRs (six waves per SIMD): the kernel is bound by arithmetic and load latency, and
// the extra waves hide more of the latter than the few spilled values cost.
template <class CT, bool SH, bool LINES>
__global__ __launch_bounds__(SURF_BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_build_surfaces(MapPtrs m, SurfaceArgs a, const ViewsArg views,
	const uint32_t* __restrict__ counters)
{
	if(counters[COUNTER_ERROR] & ERROR_BAD_SURFACE)
		return; // the host reports the invalid descriptor; nothing is written
	build_surfaces_block<CT, SH, LINES>(blockIdx.x, m, a, views);
}

// =============================================================================================
// Draw setup: edge walk -> parts -> span records
// =============================================================================================

struct Side
{
	int32_t x_start, dx_dy;
};

// Shared per-draw data of the polygon modes (rasterizer.rs:287-310 / :472-495).
struct PolyEq
{
	int64_t d_inv_z_dx, d_inv_z_dy, inv_z_k;
	int64_t d_tc_dx[2], d_tc_dy[2], tc_k[2];
};

SW_DEV PolyEq make_poly_eq(const sw_draw& d)
{
	const float* de = d.depth_equation;
	const float* te = d.tc_equation;
	PolyEq e;
	e.d_inv_z_dx = f32_to_i64_trunc_x86(INV_Z_SCALE * de[0]);
	e.d_inv_z_dy = f32_to_i64_trunc_x86(INV_Z_SCALE * de[1]);
	e.inv_z_k = f32_to_i64_trunc_x86(INV_Z_SCALE * (de[2] + (de[0] + de[1]) * 0.5f));
	e.d_tc_dx[0] = f32_to_i64_trunc_x86(TC_SCALE * te[0]);
	e.d_tc_dx[1] = f32_to_i64_trunc_x86(TC_SCALE * te[1]);
	e.d_tc_dy[0] = f32_to_i64_trunc_x86(TC_SCALE * te[2]);
	e.d_tc_dy[1] = f32_to_i64_trunc_x86(TC_SCALE * te[3]);
	e.tc_k[0] = f32_to_i64_trunc_x86(TC_SCALE * (te[4] + (te[0] + te[2]) * 0.5f));
	e.tc_k[1] = f32_to_i64_trunc_x86(TC_SCALE * (te[5] + (te[1] + te[3]) * 0.5f));
	return e;
}

SW_DEV int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }
SW_DEV int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
SW_DEV int32_t imax32(int32_t a, int32_t b) { return a > b ? a : b; }
SW_DEV int32_t imin32(int32_t a, int32_t b) { return a < b ? a : b; }

// Span of the perspective mode for one row (rasterizer.rs:327-422).
SW_DEV void span_perspective(
	const PolyEq& e, const int32_t* tsize, int32_t y, int32_t x_start_int, int32_t x_end_int, SpanRec& r)
{
	const int64_t line_inv_z = int64_t(y) * e.d_inv_z_dy + e.inv_z_k;
	const int64_t line_tc0 = int64_t(y) * e.d_tc_dy[0] + e.tc_k[0];
	const int64_t line_tc1 = int64_t(y) * e.d_tc_dy[1] + e.tc_k[1];
	const int64_t span_start_x = x_start_int;
	const int64_t span_end_x = x_end_int - 1;
	const int32_t len_m1 = (x_end_int - 1) - x_start_int;
	const int64_t inv_z_min = 1 << INV_Z_PRE_SHIFT;
	const int64_t inv_z_max = 1 << 29;
	const int32_t inv_z_start = int32_t(imax64(inv_z_min, imin64(span_start_x * e.d_inv_z_dx + line_inv_z, inv_z_max)));
	const int32_t inv_z_end = int32_t(imax64(inv_z_min, imin64(span_end_x * e.d_inv_z_dx + line_inv_z, inv_z_max)));
	const int32_t d_inv_z = len_m1 > 0 ? idiv32(inv_z_end - inv_z_start, len_m1) : 0;
	r.v[0] = inv_z_start;
	r.v[1] = d_inv_z;
	const int64_t line_tc[2] = {line_tc0, line_tc1};
#pragma unroll
	for(int i = 0; i < 2; ++i)
	{
		const int tc_max_shift = INV_Z_SHIFT - TC_SHIFT;
		const int64_t tc_max_max = 1 << 29;
		const int64_t inv_z_corr = int64_t(inv_z_start - (1 << INV_Z_PRE_SHIFT));
		const int64_t tc_max_start = imax64(0, imin64(((inv_z_corr * int64_t(tsize[i])) >> tc_max_shift) - 1, tc_max_max));
		const int32_t tc_start = int32_t(imax64(0, imin64(span_start_x * e.d_tc_dx[i] + line_tc[i], tc_max_start)));
		int32_t tc, dtc = 0;
		if(len_m1 > 0)
		{
			const int64_t tc_max_end = imax64(0,
				imin64((((inv_z_corr + int64_t(wmul(d_inv_z, len_m1))) * int64_t(tsize[i])) >> tc_max_shift) - 1, tc_max_max));
			const int32_t tc_end = int32_t(imax64(0, imin64(span_end_x * e.d_tc_dx[i] + line_tc[i], tc_max_end)));
			dtc = idiv32(tc_end - tc_start, len_m1);
			tc = tc_end > tc_start ? tc_start : tc_end - wmul(dtc, len_m1);
		}
		else
			tc = tc_start;
		r.v[2 + i * 2] = tc;
		r.v[3 + i * 2] = dtc;
	}
}

// Span of the line-z-corrected mode for one row (rasterizer.rs:528-556).
SW_DEV void span_line_z(const PolyEq& e, const int32_t* tsize, int32_t y, int32_t x_start_int, int32_t x_end_int, SpanRec& r)
{
	const int LINE_TC_SHIFT = 16;
	const int64_t line_inv_z = int64_t(y) * e.d_inv_z_dy + e.inv_z_k;
	const int64_t span_start_x = x_start_int;
	const int64_t span_end_x = x_end_int - 1;
	const int32_t len_m1 = (x_end_int - 1) - x_start_int;
	const int64_t inv_z_min = 1 << INV_Z_PRE_SHIFT;
	const int64_t inv_z_max = 1 << 29;
	const uint32_t inv_z_start = uint32_t(imin64(imax64(span_start_x * e.d_inv_z_dx + line_inv_z, inv_z_min), inv_z_max));
	const uint32_t inv_z_end = uint32_t(imin64(imax64(span_end_x * e.d_inv_z_dx + line_inv_z, inv_z_min), inv_z_max));
	const uint64_t z_start = (1u << Z_CALC_SHIFT) / (inv_z_start >> INV_Z_PRE_SHIFT);
	const uint64_t z_end = (1u << Z_CALC_SHIFT) / (inv_z_end >> INV_Z_PRE_SHIFT);
#pragma unroll
	for(int i = 0; i < 2; ++i)
	{
		const int64_t line_tc = int64_t(y) * e.d_tc_dy[i] + e.tc_k[i];
		const uint64_t tc_s = uint64_t(imax64(line_tc + span_start_x * e.d_tc_dx[i], 0));
		const uint64_t tc_e = uint64_t(imax64(line_tc + span_end_x * e.d_tc_dx[i], 0));
		const uint64_t max_tc = (uint64_t(int64_t(tsize[i])) << LINE_TC_SHIFT) - 1;
		const uint64_t a = (z_start * tc_s) >> (TC_FINAL_SHIFT - LINE_TC_SHIFT);
		const uint64_t b = (z_end * tc_e) >> (TC_FINAL_SHIFT - LINE_TC_SHIFT);
		const int32_t tzs = int32_t(a < max_tc ? a : max_tc);
		const int32_t tze = int32_t(b < max_tc ? b : max_tc);
		r.v[i * 2] = tzs;
		r.v[i * 2 + 1] = len_m1 > 0 ? idiv32(tze - tzs, len_m1) : 0;
	}
}

struct AffinePart
{
	int32_t tc_left[2], tc_right[2], d_tc_left[2], d_tc_right[2];
};

// Per-part setup of the affine mode (rasterizer.rs:597-682), including the y_start_delta step.
SW_DEV AffinePart affine_part_setup(const sw_draw& d, int32_t y_start, int32_t y_end, Side ls, Side rs, int32_t y_start_delta)
{
	const float* de = d.depth_equation;
	const float* te = d.tc_equation;
	const float y_start_f32 = fixed16_to_f32(y_start);
	const float y_end_f32 = fixed16_to_f32(y_end);
	const int32_t y_delta = y_end - y_start;
	const float x_start_left = fixed16_to_f32(ls.x_start);
	const float x_end_left = fixed16_to_f32(ls.x_start + fixed16_mul(y_delta, ls.dx_dy));
	const float x_start_right = fixed16_to_f32(rs.x_start);
	const float x_end_right = fixed16_to_f32(rs.x_start + fixed16_mul(y_delta, rs.dx_dy));
	const float z_start_left = 1.0f / (de[0] * x_start_left + de[1] * y_start_f32 + de[2]);
	const float z_end_left = 1.0f / (de[0] * x_end_left + de[1] * y_end_f32 + de[2]);
	const float z_start_right = 1.0f / (de[0] * x_start_right + de[1] * y_start_f32 + de[2]);
	const float z_end_right = 1.0f / (de[0] * x_end_right + de[1] * y_end_f32 + de[2]);
	const int32_t y_div = imax32(y_delta, FIXED16_ONE);
	AffinePart p;
#pragma unroll
	for(int i = 0; i < 2; ++i)
	{
		const int32_t max_tc = int_to_fixed16(d.texture_size[i]) - 1;
		const int32_t tsl = imin32(imax32(f32_to_fixed16_unchecked(z_start_left * (x_start_left * te[i] + y_start_f32 * te[2 + i] + te[4 + i])), 0), max_tc);
		const int32_t tel = imin32(imax32(f32_to_fixed16_unchecked(z_end_left * (x_end_left * te[i] + y_end_f32 * te[2 + i] + te[4 + i])), 0), max_tc);
		const int32_t tsr = imin32(imax32(f32_to_fixed16_unchecked(z_start_right * (x_start_right * te[i] + y_start_f32 * te[2 + i] + te[4 + i])), 0), max_tc);
		const int32_t ter = imin32(imax32(f32_to_fixed16_unchecked(z_end_right * (x_end_right * te[i] + y_end_f32 * te[2 + i] + te[4 + i])), 0), max_tc);
		p.d_tc_left[i] = fixed16_div(tel - tsl, y_div);
		p.d_tc_right[i] = fixed16_div(ter - tsr, y_div);
		p.tc_left[i] = tsl + fixed16_mul(y_start_delta, p.d_tc_left[i]);
		p.tc_right[i] = tsr + fixed16_mul(y_start_delta, p.d_tc_right[i]);
	}
	return p;
}

struct TriPart
{
	int32_t tc_left[2], d_tc_dy_left[2], d_tc_dx[2];
	int32_t light_left[3], d_light_left[3], d_light_dx[3];
};

// Corrections of fill_triangle_part (rasterizer.rs:1007-1192).
SW_DEV void tri_part_setup(const int32_t* tsize, int32_t y_start, int32_t y_end, Side ls, Side rs, int32_t y_start_delta,
	const int32_t tc_start_left[2], int32_t d_tc_dy_left[2], int32_t d_tc_dx[2], const int32_t light_start_left[3],
	int32_t d_light_left[3], int32_t d_light_dx[3], TriPart& out)
{
	const int32_t x_left = ls.x_start + fixed16_mul(y_start_delta, ls.dx_dy) + FIXED16_HALF;
	const int32_t x_right = rs.x_start + fixed16_mul(y_start_delta, rs.dx_dy) + FIXED16_HALF;
	const int32_t y_end_delta = y_end - (y_start + y_start_delta);
	const int32_t d_x_lower = imax32(x_right - x_left, 0);
	const int32_t x_upper_left = x_left + fixed16_mul(y_end_delta, ls.dx_dy);
	const int32_t x_upper_right = x_right + fixed16_mul(y_end_delta, rs.dx_dy);
	const int32_t d_x_upper = imax32(x_upper_right - x_upper_left, 0);

	for(int i = 0; i < 2; ++i)
	{
		const int32_t max_tc = int_to_fixed16(tsize[i]) - 1;
		int32_t tl = imin32(imax32(tc_start_left[i] + fixed16_mul(y_start_delta, d_tc_dy_left[i]), 0), max_tc);
		const int32_t tld = fixed16_mul(d_tc_dx[i], d_x_lower);
		if(tld > max_tc)
		{
			tl = 0;
			d_tc_dx[i] = fixed16_div(max_tc, d_x_lower);
		}
		else if(tld < -max_tc)
		{
			tl = max_tc;
			d_tc_dx[i] = fixed16_div(-max_tc, d_x_lower);
		}
		else
		{
			const int32_t tlr = tl + tld;
			if(tlr > max_tc)
				tl -= tlr - max_tc;
			else if(tlr < 0)
				tl -= tlr;
		}
		int32_t tul = imin32(imax32(tl + fixed16_mul(y_end_delta, d_tc_dy_left[i]), 0), max_tc);
		const int32_t tud = fixed16_mul(d_tc_dx[i], d_x_upper);
		if(tud > max_tc)
		{
			tul = 0;
			d_tc_dx[i] = fixed16_div(max_tc, d_x_upper);
		}
		else if(tud < -max_tc)
		{
			tul = max_tc;
			d_tc_dx[i] = fixed16_div(-max_tc, d_x_upper);
		}
		else
		{
			const int32_t tur = tul + tud;
			if(tur > max_tc)
				tul -= tur - max_tc;
			else if(tur < 0)
				tul -= tur;
		}
		d_tc_dy_left[i] = y_end_delta > 0 ? fixed16_div(tul - tl, y_end_delta) : 0;
		const int32_t tul2 = tl + fixed16_mul(y_end_delta, d_tc_dy_left[i]);
		const int32_t tur2 = tul2 + fixed16_mul(d_tc_dx[i], d_x_upper);
		if(tur2 < 0)
			d_tc_dx[i] = d_x_upper > 0 ? fixed16_div(-tul2, d_x_upper) : 0;
		else if(tur2 > max_tc)
			d_tc_dx[i] = d_x_upper > 0 ? fixed16_div(max_tc - tul2, d_x_upper) : 0;
		out.tc_left[i] = tl;
		out.d_tc_dy_left[i] = d_tc_dy_left[i];
		out.d_tc_dx[i] = d_tc_dx[i];
	}
	for(int i = 0; i < 3; ++i)
	{
		int32_t ll = imax32(light_start_left[i] + fixed16_mul(y_start_delta, d_light_left[i]), 0);
		const int32_t llr = ll + fixed16_mul(d_x_lower, d_light_dx[i]);
		if(llr < 0)
			ll -= llr;
		int32_t lul = imax32(ll + fixed16_mul(y_end_delta, d_light_left[i]), 0);
		const int32_t lur = lul + fixed16_mul(d_x_upper, d_light_dx[i]);
		if(lur < 0)
			lul -= lur;
		d_light_left[i] = y_end_delta > 0 ? fixed16_div(lul - ll, y_end_delta) : 0;
		const int32_t lul2 = ll + fixed16_mul(y_end_delta, d_light_left[i]);
		const int32_t lur2 = lul2 + fixed16_mul(d_x_upper, d_light_dx[i]);
		if(lur2 < 0)
			d_light_dx[i] = d_x_upper > 0 ? fixed16_div(-lul2, d_x_upper) : 0;
		out.light_left[i] = ll;
		out.d_light_left[i] = d_light_left[i];
		out.d_light_dx[i] = d_light_dx[i];
	}
}

// Per-thread state of one draw's setup; the draw's `lanes` threads walk its edges identically and
// split the rows of each part among themselves.
struct SetupState
{
	uint32_t lane, lanes;
	uint32_t num_parts;
	uint64_t next_span; // next free span record of this draw
	uint64_t span_end;
	bool overflow;
	bool gap; // a part did not start on the row after the previous part's last
	int32_t row_min, row_max, x_min, x_max;
	uint64_t pixels;
	int32_t row_first, row_next; // first emitted row, row after the last emitted one
	uint32_t rows0; // rows of the first emitted part
};

// Rows a part covers after clipping (rasterizer.rs:316-317, :1000-1005).
SW_DEV uint32_t part_rows(const int32_t clip[4], int32_t y_start, int32_t y_end)
{
	const int32_t a = imax32(fixed16_round_to_int(y_start), clip[1]);
	const int32_t b = imin32(fixed16_round_to_int(y_end), clip[3]);
	return a < b ? uint32_t(b - a) : 0u;
}

// Span record of row k of a part (rasterizer.rs:316-422, :500-556, :673-703, :1000-1214); updates the
// lane's bounds and pixel count.
SW_DEV void emit_span_row(const sw_draw& d, const PolyEq& e, uint32_t kind, const int32_t clip[4], int32_t y_start_int,
	uint32_t k, int32_t x_left0, int32_t ldx, int32_t x_right0, int32_t rdx, const AffinePart* ap, const TriPart* tp,
	SpanRec* __restrict__ spans, uint64_t base, SetupState& st)
{
	const int32_t y = y_start_int + int32_t(k);
	const int32_t x_left = wadd(x_left0, wmul(int32_t(k), ldx));
	const int32_t x_right = wadd(x_right0, wmul(int32_t(k), rdx));
	const int32_t x_start_int = imax32(fixed16_floor_to_int(x_left), clip[0]);
	const int32_t x_end_int = imin32(fixed16_floor_to_int(x_right), clip[2]);
	SpanRec r;
	r.x0 = 0;
	r.x1 = 0;
	for(int i = 0; i < 6; ++i)
		r.v[i] = 0;
	if(x_start_int < x_end_int)
	{
		r.x0 = x_start_int;
		r.x1 = x_end_int;
		if(kind == 0)
			span_perspective(e, d.texture_size, y, x_start_int, x_end_int, r);
		else if(kind == 1)
			span_line_z(e, d.texture_size, y, x_start_int, x_end_int, r);
		else if(kind == 2)
		{
			const int32_t tl0 = wadd(ap->tc_left[0], wmul(int32_t(k), ap->d_tc_left[0]));
			const int32_t tl1 = wadd(ap->tc_left[1], wmul(int32_t(k), ap->d_tc_left[1]));
			const int32_t tr0 = wadd(ap->tc_right[0], wmul(int32_t(k), ap->d_tc_right[0]));
			const int32_t tr1 = wadd(ap->tc_right[1], wmul(int32_t(k), ap->d_tc_right[1]));
			const int32_t x_delta = imax32(x_right - x_left, FIXED16_ONE);
			r.v[0] = tl0;
			r.v[1] = fixed16_div(tr0 - tl0, x_delta);
			r.v[2] = tl1;
			r.v[3] = fixed16_div(tr1 - tl1, x_delta);
		}
		else
		{
			const int32_t x_start_delta = int_to_fixed16(x_start_int) + FIXED16_ONE - x_left;
			const int32_t tl0 = wadd(tp->tc_left[0], wmul(int32_t(k), tp->d_tc_dy_left[0]));
			const int32_t tl1 = wadd(tp->tc_left[1], wmul(int32_t(k), tp->d_tc_dy_left[1]));
			r.v[0] = tl0 + fixed16_mul(x_start_delta, tp->d_tc_dx[0]);
			r.v[1] = tl1 + fixed16_mul(x_start_delta, tp->d_tc_dx[1]);
#pragma unroll
			for(int i = 0; i < 3; ++i)
			{
				const int32_t ll = wadd(tp->light_left[i], wmul(int32_t(k), tp->d_light_left[i]));
				r.v[2 + i] = ll + fixed16_mul(x_start_delta, tp->d_light_dx[i]);
			}
		}
		st.x_min = imin32(st.x_min, x_start_int);
		st.x_max = imax32(st.x_max, x_end_int);
		st.row_min = imin32(st.row_min, y);
		st.row_max = imax32(st.row_max, y + 1);
		st.pixels += uint64_t(x_end_int - x_start_int);
	}
	spans[base + k] = r;
}

// Emits one part: rows [y_start_int, y_end_int) spread over the draw's lanes. Bounds and pixel
// counts accumulate per thread (reduced once per draw).
// kind: 0 perspective, 1 line z, 2 affine, 3 triangle.
SW_DEV void emit_part(const sw_draw& d, const PolyEq& e, uint32_t kind, const int32_t clip[4], int32_t y_start,
	int32_t y_end, Side ls, Side rs, const AffinePart* ap, const TriPart* tp, PartRec* __restrict__ parts,
	SpanRec* __restrict__ spans, uint32_t part_slot_base, SetupState& st)
{
	const int32_t y_start_int = imax32(fixed16_round_to_int(y_start), clip[1]);
	const int32_t y_end_int = imin32(fixed16_round_to_int(y_end), clip[3]);
	if(y_start_int >= y_end_int)
		return; // no rows: the reference draws nothing for this part
	if(st.num_parts >= MAX_PARTS_PER_DRAW)
		return;
	// The parts of a draw cover consecutive rows, so the draw's span records are one run indexed
	// by row (the raster kernel relies on it; a violation is reported, never drawn).
	if(st.num_parts == 0)
	{
		st.row_first = y_start_int;
		st.rows0 = uint32_t(y_end_int - y_start_int);
	}
	else if(y_start_int != st.row_next)
		st.gap = true;
	st.row_next = y_end_int;
	const int32_t y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y_start;
	const int32_t x_left0 = ls.x_start + fixed16_mul(y_start_delta, ls.dx_dy) + FIXED16_HALF;
	const int32_t x_right0 = rs.x_start + fixed16_mul(y_start_delta, rs.dx_dy) + FIXED16_HALF;
	const uint32_t rows = uint32_t(y_end_int - y_start_int);
	const uint64_t base = st.next_span;
	const bool fits = !st.overflow && base + rows <= st.span_end;
	st.next_span += rows;

	PartRec pr;
	pr.row0 = y_start_int;
	pr.row1 = fits ? y_end_int : y_start_int; // an overflowed part draws nothing (the host re-runs)
	pr.span_base = uint32_t(base);
	pr.kind = kind;
	for(int i = 0; i < 5; ++i)
		pr.d[i] = 0;
	if(tp)
	{
		pr.d[0] = tp->d_tc_dx[0];
		pr.d[1] = tp->d_tc_dx[1];
		pr.d[2] = tp->d_light_dx[0];
		pr.d[3] = tp->d_light_dx[1];
		pr.d[4] = tp->d_light_dx[2];
	}
	if(st.lane == 0)
		parts[part_slot_base + st.num_parts] = pr;
	const uint32_t part_index = st.num_parts++;
	if(!fits)
	{
		st.overflow = true;
		return;
	}

	// Rows of the part over the draw's lanes. With several waves per draw, a part of fewer rows than
	// lanes goes to one wave (parts in turn), so that the waves work on different parts at once.
	uint32_t k0 = st.lane, kstep = st.lanes;
	if(st.lanes > 64 && rows < st.lanes)
	{
		if((part_index & (st.lanes / 64 - 1)) != (st.lane >> 6))
			return;
		k0 = st.lane & 63u;
		kstep = 64;
	}
	for(uint32_t k = k0; k < rows; k += kstep)
		emit_span_row(d, e, kind, clip, y_start_int, k, x_left0, ls.dx_dy, x_right0, rs.dx_dy, ap, tp, spans, base, st);
}

// The polygon's vertices, one per lane of the wave (lane i holds vertex i, i < num_vertices <= 24):
// the edge walk reads them with readlane instead of a chain of dependent loads.
struct LaneVerts
{
	int32_t x, y;
	SW_DEV int32_t vx(uint32_t i) const { return __builtin_amdgcn_readlane(x, i); }
	SW_DEV int32_t vy(uint32_t i) const { return __builtin_amdgcn_readlane(y, i); }
};

SW_DEV LaneVerts load_lane_verts(const int32_t* __restrict__ v, uint32_t n)
{
	const uint32_t lane = threadIdx.x & 63u;
	LaneVerts lv{0, 0};
	if(lane < n)
	{
		lv.x = v[lane * 2];
		lv.y = v[lane * 2 + 1];
	}
	return lv;
}

// The polygon edge walk of rasterizer.rs:160-270, executed identically by every lane of the wave.
// EMIT = false only counts the rows of the parts (no divisions); EMIT = true writes parts / spans.
template <bool EMIT>
SW_DEV uint32_t walk_polygon(const sw_draw& d, const LaneVerts& v, const int32_t clip[4],
	PartRec* __restrict__ parts, SpanRec* __restrict__ spans, uint32_t part_slot_base, SetupState& st)
{
	const uint32_t n = d.num_vertices;
	const uint32_t kind = d.mode == SW_TC_LINE_Z ? 1u : (d.mode == SW_TC_AFFINE ? 2u : 0u);
	PolyEq e;
	if(EMIT)
		e = make_poly_eq(d);
	uint32_t lower = 0;
	int32_t min_y = v.vy(0);
	for(uint32_t i = 0; i < n; ++i)
		if(v.vy(i) < min_y)
		{
			min_y = v.vy(i);
			lower = i;
		}
	uint32_t left = lower, right = lower;
	int32_t cur_y = min_y;
	uint32_t rows = 0;
	for(uint32_t iter = 0; iter < n; ++iter)
	{
		uint32_t nl = left + n - 1;
		if(nl >= n)
			nl -= n;
		uint32_t nr = right + 1;
		if(nr >= n)
			nr -= n;
		const int32_t lx = v.vx(left), ly = v.vy(left), nlx = v.vx(nl), nly = v.vy(nl);
		const int32_t rx = v.vx(right), ry = v.vy(right), nrx = v.vx(nr), nry = v.vy(nr);
		const int32_t dy_left = nly - ly;
		const int32_t dy_right = nry - ry;
		const int32_t next_y = imin32(nly, nry);
		const bool wide = dy_left > FIXED16_HALF && dy_right > FIXED16_HALF;
		const bool thin = !wide && dy_left > 0 && dy_right > 0 && fixed16_round_to_int(cur_y) < fixed16_round_to_int(next_y);
		if(wide || thin)
		{
			if(!EMIT)
				rows += part_rows(clip, cur_y, next_y);
			else
			{
				Side ls, rs;
				if(wide)
				{
					const int32_t dxl = fixed16_div(nlx - lx, dy_left);
					const int32_t dxr = fixed16_div(nrx - rx, dy_right);
					ls = Side{lx + fixed16_mul(dxl, cur_y - ly), dxl};
					rs = Side{rx + fixed16_mul(dxr, cur_y - ry), dxr};
				}
				else
				{
					const int32_t thin_y = int_to_fixed16(fixed16_round_to_int(cur_y)) + FIXED16_HALF;
					ls = Side{lx + fixed16_mul_div(thin_y - ly, nlx - lx, dy_left), 0};
					rs = Side{rx + fixed16_mul_div(thin_y - ry, nrx - rx, dy_right), 0};
				}
				if(kind == 2)
				{
					const int32_t y_start_int = imax32(fixed16_round_to_int(cur_y), clip[1]);
					const int32_t y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - cur_y;
					const AffinePart ap = affine_part_setup(d, cur_y, next_y, ls, rs, y_start_delta);
					emit_part(d, e, kind, clip, cur_y, next_y, ls, rs, &ap, nullptr, parts, spans, part_slot_base, st);
				}
				else
					emit_part(d, e, kind, clip, cur_y, next_y, ls, rs, nullptr, nullptr, parts, spans, part_slot_base, st);
			}
		}
		if(nl == nr)
			break;
		if(nry < nly)
			right = nr;
		else
			left = nl;
		cur_y = next_y;
	}
	return rows;
}

// Parts and spans of a polygon with the per-part arithmetic spread over lanes. The edge walk of
// walk_polygon runs first with comparisons only (identical on every lane) and keeps, in lane p, part
// p's edges, rows and span base with emit_part's bookkeeping; lane p then computes the part's sides
// (the fixed16 divisions), start positions and affine set-up and writes its PartRec; last, every
// part's rows are spread over the draw's lanes, row r of the draw's run on lane r mod lanes.
struct LanePart
{
	int32_t cur_y, next_y, y_start_int;
	uint32_t l, nl, r, nr, wide, rows, row_begin;
	uint64_t base;
	int32_t row1;
};

SW_DEV void emit_polygon(const sw_draw& d, const LaneVerts& v, const int32_t clip[4], PartRec* __restrict__ parts,
	SpanRec* __restrict__ spans, uint32_t part_slot_base, SetupState& st)
{
	const uint32_t n = d.num_vertices;
	const uint32_t kind = d.mode == SW_TC_LINE_Z ? 1u : (d.mode == SW_TC_AFFINE ? 2u : 0u);
	const uint32_t lane = threadIdx.x & 63u;
	uint32_t lower = 0;
	int32_t min_y = v.vy(0);
	for(uint32_t i = 0; i < n; ++i)
		if(v.vy(i) < min_y)
		{
			min_y = v.vy(i);
			lower = i;
		}
	uint32_t left = lower, right = lower;
	int32_t cur_y = min_y;
	LanePart mp{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
	uint32_t total_rows = 0;
	for(uint32_t iter = 0; iter < n; ++iter)
	{
		uint32_t nl = left + n - 1;
		if(nl >= n)
			nl -= n;
		uint32_t nr = right + 1;
		if(nr >= n)
			nr -= n;
		const int32_t ly = v.vy(left), nly = v.vy(nl), ry = v.vy(right), nry = v.vy(nr);
		const int32_t dy_left = nly - ly;
		const int32_t dy_right = nry - ry;
		const int32_t next_y = imin32(nly, nry);
		const bool wide = dy_left > FIXED16_HALF && dy_right > FIXED16_HALF;
		const bool thin = !wide && dy_left > 0 && dy_right > 0 && fixed16_round_to_int(cur_y) < fixed16_round_to_int(next_y);
		if(wide || thin)
		{
			const int32_t y_start_int = imax32(fixed16_round_to_int(cur_y), clip[1]);
			const int32_t y_end_int = imin32(fixed16_round_to_int(next_y), clip[3]);
			if(y_start_int < y_end_int && st.num_parts < MAX_PARTS_PER_DRAW)
			{
				if(st.num_parts == 0)
				{
					st.row_first = y_start_int;
					st.rows0 = uint32_t(y_end_int - y_start_int);
				}
				else if(y_start_int != st.row_next)
					st.gap = true;
				st.row_next = y_end_int;
				const uint32_t rows = uint32_t(y_end_int - y_start_int);
				const uint64_t base = st.next_span;
				const bool fits = !st.overflow && base + rows <= st.span_end;
				st.next_span += rows;
				if(lane == st.num_parts)
					mp = LanePart{cur_y, next_y, y_start_int, left, nl, right, nr, wide ? 1u : 0u, fits ? rows : 0u, total_rows,
						base, fits ? y_end_int : y_start_int};
				st.num_parts++;
				if(fits)
					total_rows += rows;
				else
					st.overflow = true;
			}
		}
		if(nl == nr)
			break;
		if(nry < nly)
			right = nr;
		else
			left = nl;
		cur_y = next_y;
	}
	const uint32_t np = st.num_parts;
	if(np == 0)
		return;

	// per-part set-up, lane p for part p (every lane takes part in the lane exchanges)
	const int32_t lx = lane_pull(v.x, mp.l), lyv = lane_pull(v.y, mp.l), nlx = lane_pull(v.x, mp.nl), nlyv = lane_pull(v.y, mp.nl);
	const int32_t rx = lane_pull(v.x, mp.r), ryv = lane_pull(v.y, mp.r), nrx = lane_pull(v.x, mp.nr), nryv = lane_pull(v.y, mp.nr);
	int32_t x_left0 = 0, x_right0 = 0, ldx = 0, rdx = 0;
	AffinePart ap{{0, 0}, {0, 0}, {0, 0}, {0, 0}};
	if(lane < np)
	{
		Side ls, rs;
		if(mp.wide)
		{
			const int32_t dxl = fixed16_div(nlx - lx, nlyv - lyv);
			const int32_t dxr = fixed16_div(nrx - rx, nryv - ryv);
			ls = Side{lx + fixed16_mul(dxl, mp.cur_y - lyv), dxl};
			rs = Side{rx + fixed16_mul(dxr, mp.cur_y - ryv), dxr};
		}
		else
		{
			const int32_t thin_y = int_to_fixed16(fixed16_round_to_int(mp.cur_y)) + FIXED16_HALF;
			ls = Side{lx + fixed16_mul_div(thin_y - lyv, nlx - lx, nlyv - lyv), 0};
			rs = Side{rx + fixed16_mul_div(thin_y - ryv, nrx - rx, nryv - ryv), 0};
		}
		const int32_t y_start_delta = int_to_fixed16(mp.y_start_int) + FIXED16_HALF - mp.cur_y;
		x_left0 = ls.x_start + fixed16_mul(y_start_delta, ls.dx_dy) + FIXED16_HALF;
		x_right0 = rs.x_start + fixed16_mul(y_start_delta, rs.dx_dy) + FIXED16_HALF;
		ldx = ls.dx_dy;
		rdx = rs.dx_dy;
		if(kind == 2)
			ap = affine_part_setup(d, mp.cur_y, mp.next_y, ls, rs, y_start_delta);
		if(st.lane < 64) // the first wave of the draw
		{
			PartRec pr;
			pr.row0 = mp.y_start_int;
			pr.row1 = mp.row1;
			pr.span_base = uint32_t(mp.base);
			pr.kind = kind;
			for(int i = 0; i < 5; ++i)
				pr.d[i] = 0;
			parts[part_slot_base + lane] = pr;
		}
	}

	// Rows: part q's row k is row row_begin(q) + k of the draw's run. Every readlane of part q's
	// values happens while all lanes are active (the compiler may reuse an inactive lane's copy of
	// a value inside a divergent region).
	const PolyEq e = make_poly_eq(d);
	const uint32_t lanes = st.lanes;
	for(uint32_t q = 0; q < np; ++q)
	{
		const uint32_t rows = uint32_t(__builtin_amdgcn_readlane(int32_t(mp.rows), q));
		if(rows == 0)
			continue; // the part overflowed (and so did every later one); uniform
		const uint32_t begin = uint32_t(__builtin_amdgcn_readlane(int32_t(mp.row_begin), q));
		const int32_t y_start_int = __builtin_amdgcn_readlane(mp.y_start_int, q);
		const uint64_t base = uint64_t(uint32_t(__builtin_amdgcn_readlane(int32_t(uint32_t(mp.base)), q))) |
			(uint64_t(uint32_t(__builtin_amdgcn_readlane(int32_t(uint32_t(mp.base >> 32)), q))) << 32);
		const int32_t xl0 = __builtin_amdgcn_readlane(x_left0, q), xr0 = __builtin_amdgcn_readlane(x_right0, q);
		const int32_t qldx = __builtin_amdgcn_readlane(ldx, q), qrdx = __builtin_amdgcn_readlane(rdx, q);
		AffinePart qap;
		if(kind == 2)
			for(int i = 0; i < 2; ++i)
			{
				qap.tc_left[i] = __builtin_amdgcn_readlane(ap.tc_left[i], q);
				qap.tc_right[i] = __builtin_amdgcn_readlane(ap.tc_right[i], q);
				qap.d_tc_left[i] = __builtin_amdgcn_readlane(ap.d_tc_left[i], q);
				qap.d_tc_right[i] = __builtin_amdgcn_readlane(ap.d_tc_right[i], q);
			}
		for(uint32_t k = (st.lane - begin) & (lanes - 1u); k < rows; k += lanes)
			emit_span_row(d, e, kind, clip, y_start_int, k, xl0, qldx, xr0, qrdx, &qap, nullptr, spans, base, st);
	}
}

// The triangle split of rasterizer.rs:765-982 (sorting, long edge, middle vertex).
struct TriSplit
{
	int32_t y0[2], y1[2];
	Side ls[2], rs[2];
	int32_t tc_start[2][2], dtc_dy[2][2], light_start[2][3], dlight_dy[2][3];
	int32_t d_tc_dx[2], d_light_dx[3];
};

SW_DEV void split_triangle(const int32_t* __restrict__ p, TriSplit& t)
{
	int32_t vx[3], vy[3], vtc[3][2], vl[3][3];
#pragma unroll
	for(int i = 0; i < 3; ++i)
	{
		vx[i] = p[i * 7 + 0];
		vy[i] = p[i * 7 + 1];
		vtc[i][0] = p[i * 7 + 2];
		vtc[i][1] = p[i * 7 + 3];
		vl[i][0] = p[i * 7 + 4];
		vl[i][1] = p[i * 7 + 5];
		vl[i][2] = p[i * 7 + 6];
	}
	int up, lw;
	if(vy[0] >= vy[1] && vy[0] >= vy[2])
	{
		up = 0;
		lw = vy[1] < vy[2] ? 1 : 2;
	}
	else if(vy[1] >= vy[0] && vy[1] >= vy[2])
	{
		up = 1;
		lw = vy[0] < vy[2] ? 0 : 2;
	}
	else
	{
		up = 2;
		lw = vy[0] < vy[1] ? 0 : 1;
	}
	const int mi = 3 - up - lw;
	// select the three vertices without dynamic register indexing
	int32_t ux = 0, uy = 0, lx = 0, ly = 0, mx = 0, my = 0, utc[2] = {0, 0}, ltc[2] = {0, 0}, mtc[2] = {0, 0};
	int32_t ul[3] = {0, 0, 0}, ll[3] = {0, 0, 0}, ml[3] = {0, 0, 0};
#pragma unroll
	for(int i = 0; i < 3; ++i)
	{
		if(i == up)
		{
			ux = vx[i], uy = vy[i], utc[0] = vtc[i][0], utc[1] = vtc[i][1], ul[0] = vl[i][0], ul[1] = vl[i][1], ul[2] = vl[i][2];
		}
		if(i == lw)
		{
			lx = vx[i], ly = vy[i], ltc[0] = vtc[i][0], ltc[1] = vtc[i][1], ll[0] = vl[i][0], ll[1] = vl[i][1], ll[2] = vl[i][2];
		}
		if(i == mi)
		{
			mx = vx[i], my = vy[i], mtc[0] = vtc[i][0], mtc[1] = vtc[i][1], ml[0] = vl[i][0], ml[1] = vl[i][1], ml[2] = vl[i][2];
		}
	}
	const int32_t long_edge_dy = imax32(uy - ly, FIXED16_HALF);
	const int32_t lower_part_dy = imax32(my - ly, FIXED16_HALF);
	const int32_t upper_part_dy = imax32(uy - my, FIXED16_HALF);
	const int32_t long_edge_dx_dy = fixed16_div(ux - lx, long_edge_dy);
	const int32_t long_edge_x_in_middle = lx + fixed16_mul(long_edge_dx_dy, my - ly);
	const int32_t lower_part_dx_dy = fixed16_div(mx - lx, lower_part_dy);
	const int32_t upper_part_dx_dy = fixed16_div(ux - mx, upper_part_dy);
	int32_t le_dtc[2], le_tc_mid[2], dtc_lower[2], dtc_upper[2];
#pragma unroll
	for(int i = 0; i < 2; ++i)
	{
		le_dtc[i] = fixed16_div(utc[i] - ltc[i], long_edge_dy);
		le_tc_mid[i] = ltc[i] + fixed16_mul(le_dtc[i], lower_part_dy);
		dtc_lower[i] = fixed16_div(mtc[i] - ltc[i], lower_part_dy);
		dtc_upper[i] = fixed16_div(utc[i] - mtc[i], upper_part_dy);
	}
	int32_t le_dl[3], le_l_mid[3], dl_lower[3], dl_upper[3];
#pragma unroll
	for(int i = 0; i < 3; ++i)
	{
		le_dl[i] = fixed16_div(ul[i] - ll[i], long_edge_dy);
		le_l_mid[i] = ll[i] + fixed16_mul(le_dl[i], lower_part_dy);
		dl_lower[i] = fixed16_div(ml[i] - ll[i], lower_part_dy);
		dl_upper[i] = fixed16_div(ul[i] - ml[i], upper_part_dy);
	}
	t.y0[0] = ly;
	t.y1[0] = my;
	t.y0[1] = my;
	t.y1[1] = uy;
	if(long_edge_x_in_middle >= mx)
	{
		const int32_t middle_dx = imax32(long_edge_x_in_middle - mx, FIXED16_HALF);
#pragma unroll
		for(int i = 0; i < 2; ++i)
			t.d_tc_dx[i] = fixed16_div(le_tc_mid[i] - mtc[i], middle_dx);
#pragma unroll
		for(int i = 0; i < 3; ++i)
			t.d_light_dx[i] = fixed16_div(le_l_mid[i] - ml[i], middle_dx);
		t.ls[0] = Side{lx, lower_part_dx_dy};
		t.rs[0] = Side{lx, long_edge_dx_dy};
		t.ls[1] = Side{mx, upper_part_dx_dy};
		t.rs[1] = Side{long_edge_x_in_middle, long_edge_dx_dy};
#pragma unroll
		for(int i = 0; i < 2; ++i)
		{
			t.tc_start[0][i] = ltc[i];
			t.dtc_dy[0][i] = dtc_lower[i];
			t.tc_start[1][i] = mtc[i];
			t.dtc_dy[1][i] = dtc_upper[i];
		}
#pragma unroll
		for(int i = 0; i < 3; ++i)
		{
			t.light_start[0][i] = ll[i];
			t.dlight_dy[0][i] = dl_lower[i];
			t.light_start[1][i] = ml[i];
			t.dlight_dy[1][i] = dl_upper[i];
		}
	}
	else
	{
		const int32_t middle_dx = imax32(mx - long_edge_x_in_middle, FIXED16_HALF);
#pragma unroll
		for(int i = 0; i < 2; ++i)
			t.d_tc_dx[i] = fixed16_div(mtc[i] - le_tc_mid[i], middle_dx);
#pragma unroll
		for(int i = 0; i < 3; ++i)
			t.d_light_dx[i] = fixed16_div(ml[i] - le_l_mid[i], middle_dx);
		t.ls[0] = Side{lx, long_edge_dx_dy};
		t.rs[0] = Side{lx, lower_part_dx_dy};
		t.ls[1] = Side{long_edge_x_in_middle, long_edge_dx_dy};
		t.rs[1] = Side{mx, upper_part_dx_dy};
#pragma unroll
		for(int i = 0; i < 2; ++i)
		{
			t.tc_start[0][i] = ltc[i];
			t.dtc_dy[0][i] = le_dtc[i];
			t.tc_start[1][i] = le_tc_mid[i];
			t.dtc_dy[1][i] = le_dtc[i];
		}
#pragma unroll
		for(int i = 0; i < 3; ++i)
		{
			t.light_start[0][i] = ll[i];
			t.dlight_dy[0][i] = le_dl[i];
			t.light_start[1][i] = le_l_mid[i];
			t.dlight_dy[1][i] = le_dl[i];
		}
	}
}

SW_DEV uint32_t triangle_rows(const int32_t* __restrict__ p, const int32_t clip[4])
{
	const int32_t y0 = p[1], y1 = p[8], y2 = p[15];
	// lower -> middle and middle -> upper, sorted like rasterizer.rs:775-790
	int32_t lo, mid, up;
	if(y0 >= y1 && y0 >= y2)
	{
		up = y0;
		lo = y1 < y2 ? y1 : y2;
		mid = y1 < y2 ? y2 : y1;
	}
	else if(y1 >= y0 && y1 >= y2)
	{
		up = y1;
		lo = y0 < y2 ? y0 : y2;
		mid = y0 < y2 ? y2 : y0;
	}
	else
	{
		up = y2;
		lo = y0 < y1 ? y0 : y1;
		mid = y0 < y1 ? y1 : y0;
	}
	return part_rows(clip, lo, mid) + part_rows(clip, mid, up);
}

SW_DEV void emit_triangle(const sw_draw& d, const int32_t* __restrict__ p, const int32_t clip[4], PartRec* __restrict__ parts,
	SpanRec* __restrict__ spans, uint32_t part_slot_base, SetupState& st)
{
	TriSplit t;
	split_triangle(p, t);
	const PolyEq e = PolyEq{};
#pragma unroll
	for(int pi = 0; pi < 2; ++pi)
	{
		const int32_t y_start_int = imax32(fixed16_round_to_int(t.y0[pi]), clip[1]);
		const int32_t y_end_int = imin32(fixed16_round_to_int(t.y1[pi]), clip[3]);
		if(y_start_int >= y_end_int)
			continue;
		const int32_t y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - t.y0[pi];
		int32_t tc_start[2] = {t.tc_start[pi][0], t.tc_start[pi][1]};
		int32_t dty[2] = {t.dtc_dy[pi][0], t.dtc_dy[pi][1]};
		int32_t dtx[2] = {t.d_tc_dx[0], t.d_tc_dx[1]};
		int32_t l_start[3] = {t.light_start[pi][0], t.light_start[pi][1], t.light_start[pi][2]};
		int32_t dl[3] = {t.dlight_dy[pi][0], t.dlight_dy[pi][1], t.dlight_dy[pi][2]};
		int32_t dlx[3] = {t.d_light_dx[0], t.d_light_dx[1], t.d_light_dx[2]};
		TriPart tp;
		tri_part_setup(d.texture_size, t.y0[pi], t.y1[pi], t.ls[pi], t.rs[pi], y_start_delta, tc_start, dty, dtx, l_start, dl,
			dlx, tp);
		emit_part(d, e, 3u, clip, t.y0[pi], t.y1[pi], t.ls[pi], t.rs[pi], nullptr, &tp, parts, spans, part_slot_base, st);
	}
}

SW_DEV int32_t wave_min(int32_t v)
{
	for(int off = 32; off > 0; off >>= 1)
		v = imin32(v, __shfl_xor(v, off));
	return v;
}
SW_DEV int32_t wave_max(int32_t v)
{
	for(int off = 32; off > 0; off >>= 1)
		v = imax32(v, __shfl_xor(v, off));
	return v;
}
SW_DEV uint64_t wave_sum(uint64_t v)
{
	for(int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off);
	return v;
}


// One block of draw setup. LANES threads per draw (64: a wave per draw, for batches of many draws;
// 256: a block per draw, for frames of few, large draws).
template <int LANES>
SW_DEV void setup_draws_block(uint32_t block, const SetupArgs& a, const ViewsArg& views)
{
	const sw_draw* __restrict__ draws = a.draws;
	const uint32_t num_draws = a.num_draws;
	const int32_t* __restrict__ vdata = a.vdata;
	const uint64_t num_vertex_ints = a.num_vertex_ints;
	PartRec* __restrict__ parts = a.parts;
	SpanRec* __restrict__ spans = a.spans;
	DrawRec* __restrict__ drec = a.drec;
	uint32_t* __restrict__ counters = a.counters;
	const uint64_t span_capacity = a.span_capacity, arena_texels = a.arena_texels, skybox_texels = a.skybox_texels,
				   lite_texels = a.lite_texels, portal_texels = a.portal_texels;
	constexpr uint32_t GROUPS = 256 / LANES; // draws per block
	__shared__ uint32_t group_rows[GROUPS];
	__shared__ unsigned long long block_base;
	__shared__ int32_t red_i[4][4];
	__shared__ unsigned long long red_p[4];
	// the group is uniform over each wave (LANES >= 64): readfirstlane tells the compiler, so the
	// draw's parameters load through the scalar unit and its uniform values stay in SGPRs
	const uint32_t group = __builtin_amdgcn_readfirstlane(threadIdx.x / LANES), lane = threadIdx.x % LANES;
	const uint32_t di = block * GROUPS + group;
	const bool active = di < num_draws;

	sw_draw d;
	int32_t clip[4] = {0, 0, 0, 0};
	bool valid = false;
	const int32_t* v = vdata;
	if(active)
	{
		d = draws[di];
		uint32_t lo = 0, hi = views.count;
		while(hi - lo > 1)
		{
			const uint32_t mid = (lo + hi) >> 1;
			if(views[mid].v.first_draw <= di)
				lo = mid;
			else
				hi = mid;
		}
		const sw_view& view = views[lo].v;
		// the draw's clip rect, kept inside the view (memory safety; the reference's rects always are)
		clip[0] = imax32(d.clip_rect[0], 0);
		clip[1] = imax32(d.clip_rect[1], 0);
		clip[2] = imin32(d.clip_rect[2], int32_t(view.width));
		clip[3] = imin32(d.clip_rect[3], int32_t(view.height));
		// Texture and vertex ranges must lie inside their buffers (the reference indexes them
		// unchecked in release builds, fast_math.rs:750-777; here an invalid draw is skipped).
		const uint64_t tex_texels = uint64_t(uint32_t(d.texture_size[0])) * uint32_t(d.texture_size[1]);
		valid = d.texture_size[0] > 0 && d.texture_size[1] > 0 && d.texture_size[0] <= 65536 && d.texture_size[1] <= 65536 &&
			tex_texels <= (1ull << 29) && d.blend <= SW_BLEND_ALPHA_BLEND;
		if(d.kind == SW_DRAW_POLYGON)
		{
			const uint64_t limit = d.texture_source == SW_TEXSRC_SKYBOX ? skybox_texels
				: d.texture_source == SW_TEXSRC_PORTAL                 ? portal_texels
																	   : views[lo].arena_end;
			const uint64_t tbase = d.texture_source == SW_TEXSRC_SURFACE ? view.surface_arena_offset : 0;
			valid = valid && d.texture_source <= SW_TEXSRC_PORTAL && d.num_vertices >= 3 && d.num_vertices <= SW_MAX_POLYGON_VERTICES &&
				uint64_t(d.first_vertex_int) + 2ull * d.num_vertices <= num_vertex_ints && tbase + d.texture_offset + tex_texels <= limit;
		}
		else
			valid = valid && d.kind == SW_DRAW_TRIANGLE && uint64_t(d.first_vertex_int) + 21ull <= num_vertex_ints &&
				d.texture_offset + tex_texels <= lite_texels;
		if(lane == 0)
		{
			// the texture part of the draw's record, written now so that nothing of it stays live
			// through the edge walk (the rest is written at the end)
			DrawRec& w = drec[di];
			w.tex_w = uint32_t(d.texture_size[0]);
			w.tex_h = uint32_t(d.texture_size[1]);
			w.pad = 0;
			w.tex_offset = (d.kind == SW_DRAW_POLYGON && d.texture_source == SW_TEXSRC_SURFACE ? view.surface_arena_offset : 0) +
				d.texture_offset;
		}
		v = vdata + d.first_vertex_int;
	}

	// a polygon's vertices, one per lane (a triangle's 21 values load at fixed offsets)
	LaneVerts lverts{0, 0};
	if(active && valid && d.kind == SW_DRAW_POLYGON)
		lverts = load_lane_verts(v, d.num_vertices);

	SetupState st;
	st.lane = lane;
	st.lanes = LANES;
	st.num_parts = 0;
	st.overflow = false;
	st.gap = false;
	st.row_first = 0;
	st.row_next = 0;
	st.rows0 = 0;
	st.row_min = INT32_MAX;
	st.row_max = INT32_MIN;
	st.x_min = INT32_MAX;
	st.x_max = INT32_MIN;
	st.pixels = 0;
	const uint32_t part_slot_base = di * MAX_PARTS_PER_DRAW;

	uint32_t rows = 0;
	uint64_t base = 0;
	if(a.span_stride)
	{
		// static span storage: a draw's rows never exceed its clip rect's height <= the stride
		base = uint64_t(di) * a.span_stride;
		st.next_span = base;
		st.overflow = false;
		st.span_end = base + a.span_stride;
	}
	else
	{
		// pass 1: rows of every part (no divisions), one span allocation per block
		if(active && valid)
			rows = d.kind == SW_DRAW_POLYGON ? walk_polygon<false>(d, lverts, clip, parts, spans, part_slot_base, st)
											 : triangle_rows(v, clip);
		if(lane == 0)
			group_rows[group] = rows;
		__syncthreads();
		if(threadIdx.x == 0)
		{
			unsigned long long total = 0;
			for(uint32_t g = 0; g < GROUPS; ++g)
				total += group_rows[g];
			block_base = total ? atomicAdd(reinterpret_cast<unsigned long long*>(&counters[COUNTER_SPANS_LO]), total) : 0ull;
		}
		__syncthreads();
		base = block_base;
		for(uint32_t g = 0; g < group; ++g)
			base += group_rows[g];
		st.next_span = base;
		st.overflow = base + rows > span_capacity;
		st.span_end = st.overflow ? 0 : base + rows;
		if(active && st.overflow && lane == 0 && rows)
			raise_error(counters, a.host_err, ERROR_SPAN_OVERFLOW);
	}

	// pass 2: parts and spans
	if(active && valid && !st.overflow)
	{
		if(d.kind == SW_DRAW_POLYGON)
			emit_polygon(d, lverts, clip, parts, spans, part_slot_base, st);
		else
			emit_triangle(d, v, clip, parts, spans, part_slot_base, st);
	}

	// reduce the draw's bounds and pixel count over its lanes
	int32_t x_min = wave_min(st.x_min), x_max = wave_max(st.x_max);
	int32_t row_min = wave_min(st.row_min), row_max = wave_max(st.row_max);
	uint64_t pixels = wave_sum(st.pixels);
	if(LANES > 64)
	{
		const uint32_t wave = threadIdx.x >> 6;
		if((threadIdx.x & 63) == 0)
		{
			red_i[wave][0] = x_min;
			red_i[wave][1] = x_max;
			red_i[wave][2] = row_min;
			red_i[wave][3] = row_max;
			red_p[wave] = pixels;
		}
		__syncthreads();
		for(uint32_t w = 0; w < 4; ++w)
		{
			x_min = imin32(x_min, red_i[w][0]);
			x_max = imax32(x_max, red_i[w][1]);
			row_min = imin32(row_min, red_i[w][2]);
			row_max = imax32(row_max, red_i[w][3]);
		}
		pixels = red_p[0] + red_p[1] + red_p[2] + red_p[3];
	}
	if(active && lane == 0)
	{
		const bool reads = d.blend == SW_BLEND_AVERAGE || d.blend == SW_BLEND_ADDITIVE || d.blend == SW_BLEND_ALPHA_TEST ||
			d.blend == SW_BLEND_ALPHA_BLEND;
		if(st.gap)
			raise_error(counters, a.host_err, ERROR_PART_GAP);
		const uint32_t kind = d.kind == SW_DRAW_TRIANGLE ? 3u : (d.mode == SW_TC_LINE_Z ? 1u : (d.mode == SW_TC_AFFINE ? 2u : 0u));
		DrawRec& r = drec[di];
		r.num_parts = (st.overflow || st.gap) ? 0 : st.num_parts;
		r.span_base = uint32_t(base);
		r.row_first = st.row_first;
		r.rows0 = st.rows0;
		r.row_min = row_min;
		r.row_max = row_max;
		r.x_min = x_min;
		r.x_max = x_max;
		r.pixels = uint32_t(pixels);
		r.stat = (d.kind == SW_DRAW_TRIANGLE ? DRAW_STAT_TRIANGLE : 0u) | (reads ? DRAW_STAT_READS_DST : 0u) |
			(d.kind == SW_DRAW_POLYGON && d.texture_source == SW_TEXSRC_SKYBOX ? DRAW_STAT_SKYBOX : 0u) |
			(d.kind == SW_DRAW_POLYGON && d.texture_source == SW_TEXSRC_PORTAL ? DRAW_STAT_PORTAL : 0u) |
			(kind << DRAW_STAT_KIND_SHIFT) | ((d.blend & 15u) << DRAW_STAT_BLEND_SHIFT);
		r.rows = a.span_stride ? uint32_t(st.next_span - base) : rows;
		const bool empty = r.num_parts == 0 || row_min >= row_max || x_min >= x_max;
		a.bounds[di] = empty ? DrawBounds{1, 0, 1, 0, 0, 0}
							 : DrawBounds{int16_t(x_min), int16_t(x_max), int16_t(row_min), int16_t(row_max), uint32_t(base),
								   st.row_first};
	}
}

// Draw setup; blocks past the draws' validate the surfaces and fill the surface block table
// (when the descriptors are device-resident, so the host cannot).
template <int LANES>
__global__ __launch_bounds__(256) void k_setup_draws(SetupArgs a, const ViewsArg views, const sw_surface_desc* __restrict__ surfaces,
	uint32_t num_surfaces, SurfExt* __restrict__ surf_ext, uint32_t* __restrict__ blk_surf, MapPtrs m)
{
	if(blockIdx.x >= a.setup_blocks)
	{
		const uint32_t s = (blockIdx.x - a.setup_blocks) * blockDim.x + threadIdx.x;
		if(s < num_surfaces)
			prep_surface(surfaces, s, num_surfaces, views, surf_ext, blk_surf, a.arena_texels, m, a.counters, a.host_err);
		return;
	}
	setup_draws_block<LANES>(blockIdx.x, a, views);
}

// Draw setup and surface texels in one launch (the host prepared and validated the surfaces): the
// two are independent, so they run side by side. Held to 96 VGPRs: five waves per SIMD.
template <int LANES, class CT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_setup_and_surfaces(SetupArgs a, const ViewsArg views, SurfaceArgs sa, MapPtrs m)
{
	if(blockIdx.x < a.setup_blocks)
		setup_draws_block<LANES>(blockIdx.x, a, views);
	else
		build_surfaces_block<CT, false, false>(blockIdx.x - a.setup_blocks, m, sa, views);
}

// One block after the raster kernel: sums the per-draw statistics into the result block and clears
// the live block for the next batch (a kernel boundary orders it after every other kernel's writes).
__global__ __launch_bounds__(1024) void k_finish(const DrawRec* __restrict__ drec, uint32_t num_draws, uint32_t* __restrict__ counters,
	uint32_t* __restrict__ result)
{
	unsigned long long poly = 0, tri = 0, blend = 0, parts = 0, spans = 0;
	for(uint32_t i = threadIdx.x; i < num_draws; i += blockDim.x)
	{
		const DrawRec r = drec[i];
		if(r.stat & DRAW_STAT_TRIANGLE)
			tri += r.pixels;
		else
			poly += r.pixels;
		if(r.stat & DRAW_STAT_READS_DST)
			blend += r.pixels;
		parts += r.num_parts;
		spans += r.rows;
	}
	poly = wave_sum(poly);
	tri = wave_sum(tri);
	blend = wave_sum(blend);
	parts = wave_sum(parts);
	spans = wave_sum(spans);
	__shared__ unsigned long long red[16][5];
	const uint32_t wave = threadIdx.x >> 6;
	if((threadIdx.x & 63) == 0)
	{
		red[wave][0] = poly;
		red[wave][1] = tri;
		red[wave][2] = blend;
		red[wave][3] = parts;
		red[wave][4] = spans;
	}
	__syncthreads();
	if(threadIdx.x == 0)
	{
		unsigned long long sum[5] = {0, 0, 0, 0, 0};
		for(uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w)
			for(int k = 0; k < 5; ++k)
				sum[k] += red[w][k];
		uint32_t res[COUNTER_COUNT];
		for(int k = 0; k < COUNTER_COUNT; ++k)
			res[k] = 0;
		res[COUNTER_SPANS_LO] = uint32_t(sum[4]);
		res[COUNTER_SPANS_HI] = uint32_t(sum[4] >> 32);
		res[COUNTER_POLY_PIXELS_LO] = uint32_t(sum[0]);
		res[COUNTER_POLY_PIXELS_HI] = uint32_t(sum[0] >> 32);
		res[COUNTER_TRI_PIXELS_LO] = uint32_t(sum[1]);
		res[COUNTER_TRI_PIXELS_HI] = uint32_t(sum[1] >> 32);
		res[COUNTER_BLEND_PIXELS_LO] = uint32_t(sum[2]);
		res[COUNTER_BLEND_PIXELS_HI] = uint32_t(sum[2] >> 32);
		res[COUNTER_PARTS] = uint32_t(sum[3]);
		res[COUNTER_ERROR] = counters[COUNTER_ERROR];
		for(int k = 0; k < COUNTER_COUNT; ++k)
		{
			result[k] = res[k];
			counters[k] = 0;
		}
	}
}

// =============================================================================================
// Tiled raster: one wave per 64 x ROWS tile, one pixel column per lane, colour in registers
// =============================================================================================

// Texel address of one pixel of a polygon span: perspective (rasterizer.rs:425-452) or linear tc
// (line-z :559-579, affine :705-727); -1 if the texel lies outside the texture (skipped; the
// reference's clamps keep its unchecked fetches inside).
template <int KIND>
SW_DEV int32_t polygon_texel_index(const int32_t* sv, int32_t j, uint32_t tw, uint32_t th)
{
	uint32_t u, v;
	if(KIND == 0)
	{
		const int32_t inv_z = wadd(sv[0], wmul(j, sv[1]));
		uint32_t div = uint32_t(inv_z) >> INV_Z_PRE_SHIFT;
		div = div ? div : 1u;
		// (1 << Z_CALC_SHIFT) / div, exact through the double quotient
		const uint32_t z = swdiv::z_from_inv_z(div);
		const int32_t t0 = wadd(sv[2], wmul(j, sv[3]));
		const int32_t t1 = wadd(sv[4], wmul(j, sv[5]));
		u = uint32_t((uint64_t(z) * uint64_t(int64_t(t0))) >> TC_FINAL_SHIFT);
		v = uint32_t((uint64_t(z) * uint64_t(int64_t(t1))) >> TC_FINAL_SHIFT);
	}
	else
	{
		u = uint32_t(wadd(sv[0], wmul(j, sv[1])) >> 16);
		v = uint32_t(wadd(sv[2], wmul(j, sv[3])) >> 16);
	}
	return (u < tw && v < th) ? int32_t(u + v * tw) : -1;
}

// Element i of a global-memory array whose address travels as an integer (a readlane result): the
// global address space keeps the load a global_load (a generic pointer would make it a flat load,
// which also waits on, and occupies, the LDS counter).
template <class T>
SW_DEV T global_at(uint64_t addr, int32_t i)
{
	typedef __attribute__((address_space(1))) const T* GlobalPtr;
	typedef __attribute__((address_space(1))) const char* GlobalBytes;
	// i * sizeof(T) < 2^32: textures hold at most 2^29 texels (the draw validation)
	return *(GlobalPtr)((GlobalBytes)addr + uint32_t(i) * uint32_t(sizeof(T)));
}


// Texel indices of a draw's active rows in the tile (bit rr of gmask: row rr's span meets the
// tile), the span values read from lane g * ROWS + rr: -1 outside the span or the texture. KIND 0:
// perspective, 1: linear tc, 3: triangle (upper0 = y0 - first row - rows of the lower part: row rr
// is in the upper part when upper0 + rr >= 0; its u / v steps d0 / d1 per part).
template <int KIND, int ROWS>
SW_DEV void span_texel_indices(int32_t (&tix)[ROWS], uint32_t gmask, int g, const SpanRec& mine, int32_t x, uint32_t tw,
	uint32_t th, int32_t upper0, int32_t d0u, int32_t d0l, int32_t d1u, int32_t d1l)
{
#pragma unroll
	for(int rr = 0; rr < ROWS; ++rr)
	{
		if(!(gmask & (1u << rr)))
			continue;
		const int sl = g * ROWS + rr;
		const int32_t sx0 = __builtin_amdgcn_readlane(mine.x0, sl), sx1 = __builtin_amdgcn_readlane(mine.x1, sl);
		const int32_t j = x - sx0;
		const bool in = x >= sx0 && x < sx1;
		int32_t t;
		if constexpr(KIND == 3)
		{
			const bool upper = upper0 + rr >= 0;
			const int32_t s0 = __builtin_amdgcn_readlane(mine.v[0], sl), s1 = __builtin_amdgcn_readlane(mine.v[1], sl);
			const uint32_t u = uint32_t(fixed16_floor_to_int(wadd(s0, wmul(j, upper ? d0u : d0l))));
			const uint32_t v = uint32_t(fixed16_floor_to_int(wadd(s1, wmul(j, upper ? d1u : d1l))));
			t = u < tw && v < th ? int32_t(u + v * tw) : -1;
		}
		else
		{
			int32_t sv[6];
#pragma unroll
			for(int i = 0; i < (KIND == 0 ? 6 : 4); ++i)
				sv[i] = __builtin_amdgcn_readlane(mine.v[i], sl);
			t = polygon_texel_index<KIND>(sv, j, tw, th);
		}
		tix[rr] = in ? t : -1;
	}
}

// The tile's draws are taken 64 at a time in draw order: a ballot bins them, the hits are
// compacted into the low lanes, which load their draw's parameters at once. Then GD draws at a
// time (GD * ROWS = 16 (draw, row) items): one lane per item loads its span record, every lane
// computes the texel address of its pixel for all 16 items, issues the 16 fetches together and
// blends them in painter's order into its ROWS colours.
template <class CT, int ROWS, bool SSPAN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(typename CT::T) == 4 && ROWS <= 8 ? 5 : 1)))
void k_raster_tiles(MapPtrs m, const DrawBounds* __restrict__ bounds,
	const DrawRec* __restrict__ drec, const PartRec* __restrict__ parts, const SpanRec* __restrict__ spans,
	const ViewsArg views, const typename CT::T* __restrict__ arena, const typename CT::T* __restrict__ skybox,
	typename CT::T* __restrict__ fb, uint32_t total_tiles, const uint32_t* __restrict__ counters)
{
	typedef typename CT::T C;
	constexpr int GD = ROWS >= 16 ? 1 : 16 / ROWS;
	// wave-uniform (readfirstlane): the tile, its view and every per-item value derived from them
	// stay scalar, so the uniform tests below are scalar branches
	const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
	const uint32_t t = blockIdx.x * 4 + wave;
	if(t >= total_tiles)
		return;
	// the error flag is read now and tested before the first span record is read (span records may
	// be incomplete after an overflow), so its latency overlaps the tile's first loads
	const uint32_t err = counters[COUNTER_ERROR];
	uint32_t lo = 0, hi = views.count;
	while(hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if(views[mid].first_tile <= t)
			lo = mid;
		else
			hi = mid;
	}
	const ViewExt& ve = views[lo];
	const uint32_t local = t - ve.first_tile;
	const uint32_t band = local / ve.cols;
	const uint32_t col = local - band * ve.cols;
	const int32_t x0 = int32_t(col * 64);
	const int32_t y0 = int32_t(band * ROWS);
	const int32_t W = int32_t(ve.v.width), H = int32_t(ve.v.height);
	const int32_t x = x0 + int32_t(lane);
	C* vfb = fb + ve.v.fb_offset;
	const C bg = C(ve.v.background_color);

	C acc[ROWS];
#pragma unroll
	for(int r = 0; r < ROWS; ++r)
	{
		const int32_t y = y0 + r;
		C c = bg;
		if(!ve.v.clear_background && x < W && y < H)
			c = vfb[uint64_t(y) * uint32_t(W) + uint32_t(x)];
		acc[r] = c;
	}

	const uint32_t first_draw = ve.v.first_draw, num_draws = ve.v.num_draws;
	const uint64_t lanes_below = (1ull << lane) - 1ull;
	for(uint32_t base = 0; base < num_draws; base += 64)
	{
		// ordered binning: lane i tests draw base + i against this tile
		const uint32_t di = first_draw + base + lane;
		bool hit = false;
		DrawBounds b = DrawBounds{1, 0, 1, 0, 0, 0};
		if(base + lane < num_draws)
		{
			b = bounds[di];
			hit = b.row_min < y0 + ROWS && b.row_max > y0 && b.x_min < x0 + 64 && b.x_max > x0;
		}
		const uint64_t mask = __ballot(hit);
		if(!mask)
			continue;
		const uint32_t nh = uint32_t(__popcll(mask));
		// compact the hits into lanes 0 .. nh - 1, in draw order (a permutation of the lanes); the
		// hits' full records are loaded by those lanes
		const uint32_t dst = hit ? uint32_t(__popcll(mask & lanes_below)) : nh + uint32_t(__popcll(~mask & lanes_below));
		const uint32_t h_di = uint32_t(lane_push(int32_t(di), dst));
		// With 4-row tiles in the vector span mode, the span records of the first 16 / ROWS hits are
		// loaded now, from the bounds, while the hits' full records are loaded below: lane L takes
		// draw L / ROWS at tile row L % ROWS, for a row whose span is not empty (inside the records).
		constexpr bool PREFETCH = !SSPAN && ROWS == 4;
		SpanRec next_span = SpanRec{0, 0, {0, 0, 0, 0, 0, 0}};
		bool next_ok = false;
		int32_t b_span_base = 0, b_row_first = 0, b_rmin = 1, b_rmax = 0;
		if constexpr(PREFETCH)
		{
			b_span_base = lane_push(int32_t(b.span_base), dst);
			b_row_first = lane_push(b.row_first, dst);
			b_rmin = lane_push(int32_t(b.row_min), dst);
			b_rmax = lane_push(int32_t(b.row_max), dst);
			if(err & ERROR_SPAN_OVERFLOW)
				return; // the host runs the batch again with room for every span record
		}
		auto load_group_spans_b = [&](uint32_t kgp, SpanRec& out) -> bool {
			const uint32_t kq = (kgp + lane / ROWS) & 63u;
			const int32_t y = y0 + int32_t(lane % ROWS);
			const int32_t q_base = lane_pull(b_span_base, kq);
			const int32_t q_first = lane_pull(b_row_first, kq);
			const int32_t q_min = lane_pull(b_rmin, kq), q_max = lane_pull(b_rmax, kq);
			if(lane < uint32_t(GD * ROWS) && kgp + lane / ROWS < nh && y >= q_min && y < q_max)
			{
				out = spans[uint32_t(q_base) + uint32_t(y - q_first)];
				return true;
			}
			return false;
		};
		if constexpr(PREFETCH)
			next_ok = load_group_spans_b(0, next_span);
		DrawRec r{};
		if(lane < nh)
			r = drec[h_di];
		// per-draw parameters, one hit per lane: all in the record, except a triangle's two parts'
		// steps
		const int32_t h_row_first = r.row_first;
		const int32_t h_rows = int32_t(r.rows);
		const int32_t h_span_base = int32_t(r.span_base);
		const int32_t h_rows0 = int32_t(r.rows0);
		const int32_t h_kind = int32_t((r.stat >> DRAW_STAT_KIND_SHIFT) & 3u);
		const int32_t h_blend = int32_t((r.stat >> DRAW_STAT_BLEND_SHIFT) & 15u);
		const int32_t h_tw = int32_t(r.tex_w);
		const int32_t h_th = int32_t(r.tex_h);
		const uint64_t tex = (r.stat & DRAW_STAT_TRIANGLE) ? reinterpret_cast<uint64_t>(m.lite + r.tex_offset)
			: (r.stat & DRAW_STAT_SKYBOX)                  ? reinterpret_cast<uint64_t>(skybox + r.tex_offset)
			: (r.stat & DRAW_STAT_PORTAL)                  ? reinterpret_cast<uint64_t>(static_cast<const C*>(m.portal) + r.tex_offset)
														   : reinterpret_cast<uint64_t>(arena + r.tex_offset);
		const int32_t h_tex_lo = int32_t(uint32_t(tex));
		const int32_t h_tex_hi = int32_t(uint32_t(tex >> 32));
		int32_t pd[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
		if(lane < nh && h_kind == 3)
		{
			const PartRec& p0 = parts[h_di * MAX_PARTS_PER_DRAW];
			const PartRec& p1 = parts[h_di * MAX_PARTS_PER_DRAW + 1];
#pragma unroll
			for(int i = 0; i < 5; ++i)
			{
				pd[i] = p0.d[i];
				pd[5 + i] = p1.d[i];
			}
		}

		if(err & ERROR_SPAN_OVERFLOW)
			return; // the host runs the batch again with room for every span record
		// Lane L < GD * ROWS loads the span record of draw kg + L / ROWS at tile row L % ROWS
		// (non-SSPAN mode). With 4-row tiles the next group's records are loaded while the current
		// group's texels are fetched, so each group after the first waits for one memory round
		// trip, not two (taller tiles have no registers to spare for the second set of records).
		auto load_group_spans = [&](uint32_t kgp, SpanRec& out) -> bool {
			const uint32_t kq = (kgp + lane / ROWS) & 63u;
			const int32_t q_row = y0 + int32_t(lane % ROWS) - lane_pull(h_row_first, kq);
			const int32_t q_rows = lane_pull(h_rows, kq);
			const int32_t q_base = lane_pull(h_span_base, kq);
			if(lane < uint32_t(GD * ROWS) && kgp + lane / ROWS < nh && q_row >= 0 && q_row < q_rows)
			{
				out = spans[uint32_t(q_base) + uint32_t(q_row)];
				return true;
			}
			return false;
		};
		for(uint32_t kg = 0; kg < nh; kg += GD)
		{
			if constexpr(!SSPAN)
			{
				// Each lane tests whether its record's span meets the tile; the ballot of the tests
				// names the active items, and only those are walked (scalar bit tests), their
				// values read from the item's lane. Texels of the lanes outside a span are fetched
				// at index 0 of the same texture (in bounds) and discarded by a select, so the
				// fetches and blends change no execution mask.
				SpanRec mine = SpanRec{0, 0, {0, 0, 0, 0, 0, 0}};
				bool act = false;
				if constexpr(PREFETCH)
				{
					if(next_ok)
						mine = next_span;
					act = next_ok;
					next_ok = false;
					if(kg + GD < nh)
						next_ok = load_group_spans_b(kg + GD, next_span);
				}
				else
					act = load_group_spans(kg, mine);
				act = act && mine.x0 < mine.x1 && mine.x1 > x0 && mine.x0 < x0 + 64;
				const uint64_t amask = __ballot(act);
				if(amask == 0)
					continue;
				int32_t tix[GD][ROWS];
				C texel[GD][ROWS];
#pragma unroll
				for(int g = 0; g < GD; ++g)
				{
					const uint32_t gmask = uint32_t(amask >> (g * ROWS)) & ((1u << ROWS) - 1u);
#pragma unroll
					for(int rr = 0; rr < ROWS; ++rr)
					{
						tix[g][rr] = -1;
						texel[g][rr] = 0;
					}
					if(gmask == 0)
						continue;
					const uint32_t k = kg + uint32_t(g); // below nh: the draw has an active item
					const int32_t kind = rl(h_kind, k);
					const uint32_t tw = uint32_t(rl(h_tw, k)), th = uint32_t(rl(h_th, k));
					// the draw's kind is tested once, the rows below are straight-line code
					if(kind == 0)
						span_texel_indices<0, ROWS>(tix[g], gmask, g, mine, x, tw, th, 0, 0, 0, 0, 0);
					else if(kind == 3)
						span_texel_indices<3, ROWS>(tix[g], gmask, g, mine, x, tw, th, y0 - rl(h_row_first, k) - rl(h_rows0, k),
							rl(pd[5], k), rl(pd[0], k), rl(pd[6], k), rl(pd[1], k));
					else
						span_texel_indices<1, ROWS>(tix[g], gmask, g, mine, x, tw, th, 0, 0, 0, 0, 0);
				}
				// every span value is read before the fetches are issued, so all of the group's
				// fetches are in flight together
#pragma unroll
				for(int g = 0; g < GD; ++g)
				{
					const uint32_t gmask = uint32_t(amask >> (g * ROWS)) & ((1u << ROWS) - 1u);
					if(gmask == 0)
						continue;
					const uint32_t k = kg + uint32_t(g);
					const uint64_t tex = uint64_t(uint32_t(rl(h_tex_lo, k))) | (uint64_t(uint32_t(rl(h_tex_hi, k))) << 32);
					if(rl(h_kind, k) == 3)
					{
#pragma unroll
						for(int rr = 0; rr < ROWS; ++rr)
							if(gmask & (1u << rr))
								texel[g][rr] = C(global_at<uint32_t>(tex, tix[g][rr] < 0 ? 0 : tix[g][rr]));
					}
					else
					{
#pragma unroll
						for(int rr = 0; rr < ROWS; ++rr)
							if(gmask & (1u << rr))
								texel[g][rr] = global_at<C>(tex, tix[g][rr] < 0 ? 0 : tix[g][rr]);
					}
				}
#pragma unroll
				for(int g = 0; g < GD; ++g)
				{
					const uint32_t gmask = uint32_t(amask >> (g * ROWS)) & ((1u << ROWS) - 1u);
					if(gmask == 0)
						continue;
					const uint32_t k = kg + uint32_t(g);
					const uint32_t blend = uint32_t(rl(h_blend, k));
					if(rl(h_kind, k) != 3)
					{
						if(blend == SW_BLEND_NONE)
						{
#pragma unroll
							for(int rr = 0; rr < ROWS; ++rr)
								if(gmask & (1u << rr))
									acc[rr] = tix[g][rr] >= 0 ? texel[g][rr] : acc[rr];
						}
						else
						{
#pragma unroll
							for(int rr = 0; rr < ROWS; ++rr)
							{
								if(!(gmask & (1u << rr)))
									continue;
								const C nc = write_into_framebuffer<CT>(blend, acc[rr], texel[g][rr]);
								acc[rr] = tix[g][rr] >= 0 ? nc : acc[rr];
							}
						}
						continue;
					}
					const int32_t part_row0 = rl(h_row_first, k), rows0 = rl(h_rows0, k);
					const int32_t dr_u = rl(pd[7], k), dr_l = rl(pd[2], k);
					const int32_t dg_u = rl(pd[8], k), dg_l = rl(pd[3], k);
					const int32_t db_u = rl(pd[9], k), db_l = rl(pd[4], k);
#pragma unroll
					for(int rr = 0; rr < ROWS; ++rr)
					{
						if(!(gmask & (1u << rr)))
							continue;
						const int sl = g * ROWS + rr;
						const bool upper = y0 + rr - part_row0 >= rows0;
						const int32_t j = x - rl(mine.x0, sl);
						const int32_t lr = wadd(rl(mine.v[2], sl), wmul(j, upper ? dr_u : dr_l));
						const int32_t lg = wadd(rl(mine.v[3], sl), wmul(j, upper ? dg_u : dg_l));
						const int32_t lb = wadd(rl(mine.v[4], sl), wmul(j, upper ? db_u : db_l));
						const uint32_t tv = uint32_t(texel[g][rr]);
						const CI tc = Color32T::to_ci(tv);
						const CI lit = CI{wmul(tc.b, lb) >> 16, wmul(tc.g, lg) >> 16, wmul(tc.r, lr) >> 16, 0};
						const C dst_c = acc[rr];
						C nc = dst_c;
						switch(blend)
						{
						case SW_BLEND_NONE: nc = CT::from_ci(lit); break;
						case SW_BLEND_AVERAGE: nc = CT::average(dst_c, CT::from_ci(lit)); break;
						case SW_BLEND_ADDITIVE: nc = CT::saturated_sum(dst_c, CT::from_ci(lit)); break;
						case SW_BLEND_ALPHA_TEST: nc = Color32T::test_alpha(tv) ? CT::from_ci(lit) : dst_c; break;
						case SW_BLEND_ALPHA_BLEND: nc = alpha_blend_ci<CT>(dst_c, lit, Color32T::get_alpha(tv)); break;
						default: break;
						}
						acc[rr] = tix[g][rr] >= 0 ? nc : dst_c;
					}
				}
				continue;
			}
			// Span records of the group's 16 items. SSPAN: each item's record is loaded through the
			// scalar unit into SGPRs when the item is processed. Otherwise lane L < GD * ROWS loads
			// the record of draw kg + L / ROWS at tile row L % ROWS and items read them with
			// readlane (fewer, wider loads; better where many distinct records are in flight).
			SpanRec mine = SpanRec{0, 0, {0, 0, 0, 0, 0, 0}};
			if constexpr(!SSPAN)
			{
				const uint32_t kq = (kg + lane / ROWS) & 63u;
				const int32_t q_row = y0 + int32_t(lane % ROWS) - lane_pull(h_row_first, kq);
				const int32_t q_rows = lane_pull(h_rows, kq);
				const int32_t q_base = lane_pull(h_span_base, kq);
				if(lane < uint32_t(GD * ROWS) && kg + lane / ROWS < nh && q_row >= 0 && q_row < q_rows)
					mine = spans[uint32_t(q_base) + uint32_t(q_row)];
			}
			int32_t tix[GD][ROWS];
#pragma unroll
			for(int g = 0; g < GD; ++g)
			{
				const uint32_t k = kg + uint32_t(g);
				const uint32_t kk = k < nh ? k : 0u;
				const int32_t kind = rl(h_kind, kk);
				const uint32_t tw = uint32_t(rl(h_tw, kk)), th = uint32_t(rl(h_th, kk));
				const int32_t part_row0 = rl(h_row_first, kk), rows0 = rl(h_rows0, kk);
				const int32_t nrows = rl(h_rows, kk), span_base = rl(h_span_base, kk);
#pragma unroll
				for(int rr = 0; rr < ROWS; ++rr)
				{
					// every value of the item is uniform over the wave, its span record is loaded
					// through the scalar unit straight into SGPRs, and every branch below is a
					// scalar one; the lane's own span test only masks the result
					const int32_t q = y0 + rr - part_row0;
					const int sl = g * ROWS + rr;
					int32_t ti = -1;
					SpanRec sp = SpanRec{0, 0, {0, 0, 0, 0, 0, 0}};
					bool row_ok;
					if constexpr(SSPAN)
					{
						row_ok = k < nh && q >= 0 && q < nrows;
						if(row_ok)
							sp = spans[uint32_t(span_base) + uint32_t(q)];
					}
					else
					{
						row_ok = k < nh;
						sp.x0 = rl(mine.x0, sl);
						sp.x1 = rl(mine.x1, sl);
					}
					if(row_ok)
					{
						const int32_t sx0 = sp.x0, sx1 = sp.x1;
						if(sx0 < sx1 && sx1 > x0 && sx0 < x0 + 64)
						{
							if constexpr(!SSPAN)
							{
#pragma unroll
								for(int i = 0; i < 6; ++i)
									sp.v[i] = rl(mine.v[i], sl);
							}
							const int32_t* sv = sp.v;
							const int32_t j = x - sx0;
							const bool in = x >= sx0 && x < sx1;
							if(kind == 3)
							{
								const bool upper = q >= rows0;
								const int32_t d0u = rl(pd[5], kk), d0l = rl(pd[0], kk);
								const int32_t d1u = rl(pd[6], kk), d1l = rl(pd[1], kk);
								const int32_t d0 = upper ? d0u : d0l, d1 = upper ? d1u : d1l;
								const uint32_t u = uint32_t(fixed16_floor_to_int(wadd(sv[0], wmul(j, d0))));
								const uint32_t v = uint32_t(fixed16_floor_to_int(wadd(sv[1], wmul(j, d1))));
								ti = in && u < tw && v < th ? int32_t(u + v * tw) : -1;
							}
							else if(kind == 0)
							{
								const int32_t t = polygon_texel_index<0>(sv, j, tw, th);
								ti = in ? t : -1;
							}
							else
							{
								const int32_t t = polygon_texel_index<1>(sv, j, tw, th);
								ti = in ? t : -1;
							}
						}
					}
					tix[g][rr] = ti;
				}
			}

			C texel[GD][ROWS];
#pragma unroll
			for(int g = 0; g < GD; ++g)
			{
				const uint32_t k = kg + uint32_t(g);
				const uint32_t kk = k < nh ? k : 0u;
				const bool tri = rl(h_kind, kk) == 3;
				const uint64_t tex = uint64_t(uint32_t(rl(h_tex_lo, kk))) | (uint64_t(uint32_t(rl(h_tex_hi, kk))) << 32);
#pragma unroll
				for(int rr = 0; rr < ROWS; ++rr)
				{
					C tv = 0;
					if(tix[g][rr] >= 0)
						tv = tri ? C(global_at<uint32_t>(tex, tix[g][rr])) : global_at<C>(tex, tix[g][rr]);
					texel[g][rr] = tv;
				}
			}

#pragma unroll
			for(int g = 0; g < GD; ++g)
			{
				const uint32_t k = kg + uint32_t(g);
				if(k >= nh)
					continue;
				const uint32_t blend = uint32_t(rl(h_blend, k));
				if(rl(h_kind, k) != 3)
				{
#pragma unroll
					for(int rr = 0; rr < ROWS; ++rr)
						if(tix[g][rr] >= 0)
							acc[rr] = write_into_framebuffer<CT>(blend, acc[rr], texel[g][rr]);
					continue;
				}
				const int32_t part_row0 = rl(h_row_first, k), rows0 = rl(h_rows0, k);
				const int32_t nrows = rl(h_rows, k), span_base = rl(h_span_base, k);
				const int32_t dr_u = rl(pd[7], k), dr_l = rl(pd[2], k);
				const int32_t dg_u = rl(pd[8], k), dg_l = rl(pd[3], k);
				const int32_t db_u = rl(pd[9], k), db_l = rl(pd[4], k);
#pragma unroll
				for(int rr = 0; rr < ROWS; ++rr)
				{
					const int32_t q = y0 + rr - part_row0;
					if(q < 0 || q >= nrows)
						continue; // no span in this row (so no texel either)
					const bool upper = q >= rows0;
					int32_t sx0, l0, l1, l2;
					if constexpr(SSPAN)
					{
						const SpanRec sp = spans[uint32_t(span_base) + uint32_t(q)]; // scalar load, cached
						sx0 = sp.x0;
						l0 = sp.v[2];
						l1 = sp.v[3];
						l2 = sp.v[4];
					}
					else
					{
						const int sl = g * ROWS + rr;
						sx0 = rl(mine.x0, sl);
						l0 = rl(mine.v[2], sl);
						l1 = rl(mine.v[3], sl);
						l2 = rl(mine.v[4], sl);
					}
					if(tix[g][rr] < 0)
						continue;
					const int32_t j = x - sx0;
					const int32_t dr = upper ? dr_u : dr_l;
					const int32_t dg = upper ? dg_u : dg_l;
					const int32_t db = upper ? db_u : db_l;
					const int32_t lr = wadd(l0, wmul(j, dr));
					const int32_t lg = wadd(l1, wmul(j, dg));
					const int32_t lb = wadd(l2, wmul(j, db));
					const uint32_t tv = uint32_t(texel[g][rr]);
					const CI tc = Color32T::to_ci(tv);
					const CI lit = CI{wmul(tc.b, lb) >> 16, wmul(tc.g, lg) >> 16, wmul(tc.r, lr) >> 16, 0};
					C& dst_c = acc[rr];
					switch(blend)
					{
					case SW_BLEND_NONE: dst_c = CT::from_ci(lit); break;
					case SW_BLEND_AVERAGE: dst_c = CT::average(dst_c, CT::from_ci(lit)); break;
					case SW_BLEND_ADDITIVE: dst_c = CT::saturated_sum(dst_c, CT::from_ci(lit)); break;
					case SW_BLEND_ALPHA_TEST:
						if(Color32T::test_alpha(tv))
							dst_c = CT::from_ci(lit);
						break;
					case SW_BLEND_ALPHA_BLEND: dst_c = alpha_blend_ci<CT>(dst_c, lit, Color32T::get_alpha(tv)); break;
					default: break;
					}
				}
			}
		}
	}

	if(x < W)
	{
#pragma unroll
		for(int r = 0; r < ROWS; ++r)
		{
			const int32_t y = y0 + r;
			if(y < H) // non-temporal: the finished tile is not read again by this kernel
				__builtin_nontemporal_store(acc[r], &vfb[uint64_t(y) * uint32_t(W) + uint32_t(x)]);
		}
	}
}

// =============================================================================================
// Bloom (postprocessor.rs:258-956)
// =============================================================================================

// TonemappingFunction::do_it (postprocessor.rs:1019-1023): c / fma(c, 1/255, 1/(exposure*modulate))
// per lane. The alpha lane's constants are 0, so it is c.a / +0: +inf for c.a > 0 and NaN for
// c.a == 0, and the pack's min(x, 255) (minps: the second operand when either is NaN) makes both
// 255 — the alpha byte is the constant 0xFF.
SW_DEV uint32_t tonemap_cv(CV c, const float* inv_scale)
{
	const float inv_255 = 1.0f / 255.0f;
	return lane_to_u8(c.b / mul_add(c.b, inv_255, inv_scale[0])) | (lane_to_u8(c.g / mul_add(c.g, inv_255, inv_scale[1])) << 8) |
		(lane_to_u8(c.r / mul_add(c.r, inv_255, inv_scale[2])) << 16) | 0xFF000000u;
}

SW_DEV CV cv_add(CV a, CV b) { return CV{a.b + b.b, a.g + b.g, a.r + b.r, a.a + b.a}; }
SW_DEV CI ci_add(CI a, CI b) { return CI{wadd(a.b, b.b), wadd(a.g, b.g), wadd(a.r, b.r), wadd(a.a, b.a)}; }
SW_DEV CI ci_mul(CI a, int32_t s) { return CI{wmul(a.b, s), wmul(a.g, s), wmul(a.r, s), wmul(a.a, s)}; }
SW_DEV CI ci_shr(CI a, int k) { return CI{a.b >> k, a.g >> k, a.r >> k, a.a >> k}; }

// downscale_hdr_buffer (:805-853): one block per bloom row. Each bloom texel is the integer average
// of its scale x scale source pixels ((sum * (256 / scale^2)) >> 8, packed with unsigned
// saturation); the block also sums the row's averages (integers, so any order is exact).
template <uint32_t S>
__global__ __launch_bounds__(256) void k_bloom_downscale(BloomArgs a)
{
	const uint32_t dy = blockIdx.x;
	const int32_t scaler = (1 << 8) / int32_t(S * S);
	const uint64_t* __restrict__ src = a.hdr + uint64_t(dy * S) * a.w;
	uint64_t* __restrict__ dst = a.buf0 + uint64_t(dy) * a.bw;
	CI line = CI{0, 0, 0, 0};
#pragma unroll 2
	for(uint32_t dx = threadIdx.x; dx < a.bw; dx += blockDim.x)
	{
		CI sum = CI{0, 0, 0, 0};
#pragma unroll
		for(uint32_t y = 0; y < S; ++y)
#pragma unroll
			for(uint32_t x = 0; x < S; ++x)
				sum = ci_add(sum, Color64T::to_ci(src[uint64_t(y) * a.w + dx * S + x]));
		const CI average = ci_shr(ci_mul(sum, scaler), 8);
		dst[dx] = Color64T::from_ci(average);
		line = ci_add(line, average);
	}
	__shared__ int32_t red[4][4];
	int32_t v[4] = {line.b, line.g, line.r, line.a};
#pragma unroll
	for(int k = 0; k < 4; ++k)
		for(int off = 32; off > 0; off >>= 1)
			v[k] = wadd(v[k], __shfl_xor(v[k], off));
	const uint32_t wave = threadIdx.x >> 6;
	if((threadIdx.x & 63) == 0)
		for(int k = 0; k < 4; ++k)
			red[wave][k] = v[k];
	__syncthreads();
	if(threadIdx.x < 4)
	{
		int32_t t = 0;
		for(uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w)
			t = wadd(t, red[w][threadIdx.x]);
		a.line_sums[dy * 4 + threadIdx.x] = t;
	}
}

// The average colour (:838-852): the rows' integer sums converted and added in row order, then
// scaled by 1 / (bloom texels). Float addition is not associative, so each lane is one sequential
// sum, like the reference's; the block converts chunks of rows into LDS in parallel (one 16-byte
// load per row) and four threads (one per lane) add them up in order. `a.average` is host memory.
#define BLOOM_AVG_CHUNK 512u
SW_DEV void bloom_average_block(const BloomArgs& a)
{
	__shared__ float4 rows[BLOOM_AVG_CHUNK];
	float sum = 0.0f;
	const uint32_t lane = threadIdx.x;
	const int4* line_sums = reinterpret_cast<const int4*>(a.line_sums);
	for(uint32_t y0 = 0; y0 < a.bh; y0 += BLOOM_AVG_CHUNK)
	{
		const uint32_t n = min(BLOOM_AVG_CHUNK, a.bh - y0);
		for(uint32_t i = threadIdx.x; i < n; i += blockDim.x)
		{
			const int4 l = line_sums[y0 + i];
			rows[i] = make_float4(float(l.x), float(l.y), float(l.z), float(l.w));
		}
		__syncthreads();
		if(lane < 4)
		{
			const float* r = reinterpret_cast<const float*>(rows) + lane;
#pragma unroll 8
			for(uint32_t y = 0; y < n; ++y)
				sum += r[y * 4];
		}
		__syncthreads();
	}
	if(lane < 4)
		a.average[lane] = sum * (1.0f / float(a.bw * a.bh));
}

__global__ __launch_bounds__(256) void k_bloom_average(BloomArgs a) { bloom_average_block(a); }

// perform_blur_impl (:855-937), one pass: pass 0 horizontal (buf0 -> buf1, weights k1), pass 1
// vertical (buf1 -> buf0, weights k2 = k1 scaled by bloom_scale); edge texels repeat (clamped
// coordinates); integer weights at scale 2^14, arithmetic shift, unsigned-saturating pack.
// Grid: x over the row in blocks of BLUR_PX * 256 texels (BLUR_PX per thread, their taps loaded
// together), y = bloom row. with_average: block (0, 0), dispatched first, also computes the
// average colour (pass 0 only; it reads the line sums the downscale left).
#define BLUR_PX 2
__global__ __launch_bounds__(256) void k_bloom_blur(BloomArgs a, uint32_t pass, uint32_t with_average)
{
	if(with_average && blockIdx.x == 0 && blockIdx.y == 0)
		bloom_average_block(a);
	const uint32_t y = blockIdx.y;
	const int32_t R = a.radius;
	const int32_t* k = pass == 0 ? a.k1 : a.k2;
	const uint64_t* __restrict__ src = pass == 0 ? a.buf0 : a.buf1;
	uint64_t* __restrict__ dst = pass == 0 ? a.buf1 : a.buf0;
	const int32_t last = pass == 0 ? int32_t(a.bw) - 1 : int32_t(a.bh) - 1;
	CI sum[BLUR_PX];
	uint32_t xs[BLUR_PX];
#pragma unroll
	for(int j = 0; j < BLUR_PX; ++j)
	{
		// out-of-row threads work on the row's last texel and store nothing
		xs[j] = min(blockIdx.x * (256u * BLUR_PX) + uint32_t(j) * 256u + threadIdx.x, a.bw - 1);
		const int32_t c = pass == 0 ? int32_t(xs[j]) : int32_t(y);
		const uint64_t v = pass == 0 ? src[uint64_t(y) * a.bw + uint32_t(c)] : src[uint64_t(c) * a.bw + xs[j]];
		sum[j] = ci_mul(Color64T::to_ci(v), k[R]);
	}
	for(int32_t d = 1; d <= R; ++d)
	{
		const int32_t w = k[d + R];
#pragma unroll
		for(int j = 0; j < BLUR_PX; ++j)
		{
			const int32_t c = pass == 0 ? int32_t(xs[j]) : int32_t(y);
			const int32_t p0 = imax32(c - d, 0), p1 = imin32(c + d, last);
			uint64_t v0, v1;
			if(pass == 0)
			{
				v0 = src[uint64_t(y) * a.bw + uint32_t(p0)];
				v1 = src[uint64_t(y) * a.bw + uint32_t(p1)];
			}
			else
			{
				v0 = src[uint64_t(p0) * a.bw + xs[j]];
				v1 = src[uint64_t(p1) * a.bw + xs[j]];
			}
			sum[j] = ci_add(sum[j], ci_mul(ci_add(Color64T::to_ci(v0), Color64T::to_ci(v1)), w));
		}
	}
#pragma unroll
	for(int j = 0; j < BLUR_PX; ++j)
	{
		const uint32_t x = blockIdx.x * (256u * BLUR_PX) + uint32_t(j) * 256u + threadIdx.x;
		if(x < a.bw)
			dst[uint64_t(y) * a.bw + x] = Color64T::from_ci(ci_shr(sum[j], 14));
	}
}

// perform_tonemapping_with_bloom_{nearest,linear} (:346-726), one thread per pixel: the bloom value
// of the pixel's region (block, border or corner of the reference's loops), added to the pixel,
// tonemapped.
// Grid: x over the row in blocks of TM_PX * 256 pixels (TM_PX per thread, their frame loads issued
// together), y = frame row.
#define TM_PX 4
template <bool LINEAR>
SW_DEV uint32_t bloom_pixel(const BloomArgs& a, uint32_t x, uint32_t y, CV c)
{
	const uint32_t sl = a.scale_log2, S = 1u << sl;
	auto bl = [&](uint32_t bx, uint32_t by) { return cv_from_color64(a.buf0[uint64_t(by) * a.bw + bx]); };
	CV sum;
	if(!LINEAR)
		sum = cv_add(bl(std::min(x >> sl, a.bw - 1), std::min(y >> sl, a.bh - 1)), c);
	else
	{
		const uint32_t half = S / 2;
		const float sf = float(S);
		auto mix = [&](uint32_t k) { return float((S + 1) & 1) / (2.0f * sf) + float(k) / sf; };
		const uint32_t right_x = (a.bw - 1) * S + half, upper_y = (a.bh - 1) * S + half;
		const bool lower = y < half, upper = y >= upper_y, left = x < half, right = x >= right_x;
		if((lower || upper) && (left || right))
			sum = cv_add(bl(left ? 0 : a.bw - 1, lower ? 0 : a.bh - 1), c); // corners
		else if(lower || upper)
		{
			// lower / upper border: horizontal mix of the first / last bloom row
			const uint32_t by = lower ? 0 : a.bh - 1;
			const uint32_t sx = (x - half) >> sl, dx = (x - half) & (S - 1);
			const float mx = mix(dx);
			sum = cv_mul_scalar_add(bl(sx, by), 1.0f - mx, cv_mul_scalar_add(bl(sx + 1, by), mx, c));
		}
		else
		{
			const uint32_t sy = (y - half) >> sl, dy = (y - half) & (S - 1);
			const float my = mix(dy), omy = 1.0f - my;
			if(left || right)
			{
				const uint32_t bx = left ? 0 : a.bw - 1;
				sum = cv_add(cv_mul_scalar_add(bl(bx, sy), omy, cv_scalar_mul(bl(bx, sy + 1), my)), c);
			}
			else
			{
				const uint32_t sx = (x - half) >> sl, dx = (x - half) & (S - 1);
				const float mx = mix(dx);
				const CV l = cv_mul_scalar_add(bl(sx, sy), omy, cv_scalar_mul(bl(sx, sy + 1), my));
				const CV r = cv_mul_scalar_add(bl(sx + 1, sy), omy, cv_scalar_mul(bl(sx + 1, sy + 1), my));
				sum = cv_mul_scalar_add(l, 1.0f - mx, cv_mul_scalar_add(r, mx, c));
			}
		}
	}
	return tonemap_cv(sum, a.inv_scale);
}

template <bool LINEAR>
__global__ __launch_bounds__(256) void k_bloom_tonemap(BloomArgs a)
{
	const uint32_t y = blockIdx.y;
	const uint32_t x0 = blockIdx.x * (256u * TM_PX) + threadIdx.x;
	const uint64_t* __restrict__ src = a.hdr + uint64_t(y) * a.w;
	uint64_t v[TM_PX];
#pragma unroll
	for(int j = 0; j < TM_PX; ++j)
		v[j] = src[min(x0 + uint32_t(j) * 256u, a.w - 1)];
#pragma unroll
	for(int j = 0; j < TM_PX; ++j)
	{
		const uint32_t x = x0 + uint32_t(j) * 256u;
		if(x < a.w)
			__builtin_nontemporal_store(bloom_pixel<LINEAR>(a, x, y, cv_from_color64(v[j])), &a.dst[uint64_t(y) * a.dst_pitch + x]);
	}
}

// =============================================================================================
// Tonemap (postprocessor.rs:237-253, :1019-1023)
// =============================================================================================
// Tonemap blocks walk the frame's rows (block b: rows b, b + grid, ...; the threads over the row);
// each block leaves the partial colour sums of its pixels (no atomics on three shared addresses),
// k_tonemap_sums adds them up.

__global__ __launch_bounds__(256) void k_tonemap(const uint64_t* __restrict__ hdr, uint32_t width, uint32_t height,
	float inv_scale_r, float inv_scale_g, float inv_scale_b, uint32_t* __restrict__ dst, uint32_t dst_pitch,
	float* __restrict__ partials)
{
	__shared__ float red[3][4];
	float sb = 0.0f, sg = 0.0f, sr = 0.0f;
	const float inv_scale[3] = {inv_scale_b, inv_scale_g, inv_scale_r};
	for(uint32_t y = blockIdx.x; y < height; y += gridDim.x)
	{
		const uint64_t* __restrict__ src = hdr + uint64_t(y) * width;
		uint32_t* __restrict__ out = dst + uint64_t(y) * dst_pitch;
		for(uint32_t x0 = threadIdx.x; x0 < width; x0 += 256u * TM_PX)
		{
			uint64_t v[TM_PX];
#pragma unroll
			for(int j = 0; j < TM_PX; ++j)
			{
				const uint32_t x = x0 + uint32_t(j) * 256u;
				v[j] = x < width ? src[x] : 0; // zero: adds nothing to the sums
			}
#pragma unroll
			for(int j = 0; j < TM_PX; ++j)
			{
				const uint32_t x = x0 + uint32_t(j) * 256u;
				const CV c = cv_from_color64(v[j]);
				sb += c.b;
				sg += c.g;
				sr += c.r;
				if(x < width)
					__builtin_nontemporal_store(tonemap_cv(c, inv_scale), &out[x]); // the final image: not read again here
			}
		}
	}
	for(int off = 32; off > 0; off >>= 1)
	{
		sr += __shfl_xor(sr, off);
		sg += __shfl_xor(sg, off);
		sb += __shfl_xor(sb, off);
	}
	const uint32_t wave = threadIdx.x >> 6;
	if((threadIdx.x & 63) == 0)
	{
		red[0][wave] = sr;
		red[1][wave] = sg;
		red[2][wave] = sb;
	}
	__syncthreads();
	if(threadIdx.x < 3)
		partials[blockIdx.x * 3 + threadIdx.x] =
			red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

__global__ __launch_bounds__(256) void k_tonemap_sums(const float* __restrict__ partials, uint32_t blocks, float* __restrict__ sums)
{
	__shared__ float red[3][4];
	float s[3] = {0.0f, 0.0f, 0.0f};
	for(uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
		for(int c = 0; c < 3; ++c)
			s[c] += partials[b * 3 + c];
	for(int c = 0; c < 3; ++c)
		for(int off = 32; off > 0; off >>= 1)
			s[c] += __shfl_xor(s[c], off);
	const uint32_t wave = threadIdx.x >> 6;
	if((threadIdx.x & 63) == 0)
		for(int c = 0; c < 3; ++c)
			red[c][wave] = s[c];
	__syncthreads();
	if(threadIdx.x < 3)
		sums[threadIdx.x] = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

// A portal texture's overlay (partial_renderer.rs:2205-2276 -> mix_surface_with_texture,
// surfaces.rs:903-971): block (x, r) takes texels [x * 1024, x * 1024 + 1024) of record r's rect,
// four per thread, each fetching its wrapped lite texel and blending it into the texel in place.
// The rects of one launch do not overlap (the host checks), so every texel has one writer.
constexpr uint32_t PORTAL_MIX_BLOCK_TEXELS = 1024;

template <class CT>
__global__ __launch_bounds__(256) void k_portal_mix(const uint32_t* __restrict__ lite, const swk::PortalMixRec* __restrict__ recs,
	typename CT::T* __restrict__ texels)
{
	const swk::PortalMixRec r = recs[blockIdx.y];
	const uint64_t n = uint64_t(r.w) * r.h;
	// cvtps2dq of light * 256, lanes B, G, R (surfaces.rs:911-914)
	const int32_t lb = f32_to_i32_rne_x86(r.light[2] * 256.0f);
	const int32_t lg = f32_to_i32_rne_x86(r.light[1] * 256.0f);
	const int32_t lr = f32_to_i32_rne_x86(r.light[0] * 256.0f);
#pragma unroll
	for(uint32_t k = 0; k < PORTAL_MIX_BLOCK_TEXELS / 256; ++k)
	{
		const uint64_t i = uint64_t(blockIdx.x) * PORTAL_MIX_BLOCK_TEXELS + k * 256 + threadIdx.x;
		if(i >= n)
			break;
		const uint32_t y = uint32_t(i / r.w), x = uint32_t(i - uint64_t(y) * r.w);
		// u steps from rem_euclid(tc_min[0]) with a wrap at the width: (tc_min[0] + x) mod w, exact;
		// v is the wrapping i32 sum tc_min[1] + y taken rem_euclid (surfaces.rs:920-923)
		int64_t su = (int64_t(r.tcu) + x) % int64_t(r.tw);
		int32_t sv = int32_t(uint32_t(r.tcv) + y) % int32_t(r.th);
		su += su < 0 ? r.tw : 0;
		sv += sv < 0 ? int32_t(r.th) : 0;
		const uint32_t t = lite[uint64_t(r.tex_offset) + uint64_t(sv) * r.tw + uint64_t(su)];
		typename CT::T& dst = texels[r.dst + i];
		dst = emissive_blend<CT>(t, lb, lg, lr, r.blend, dst);
	}
}

__global__ void k_zero_floats(float* p, uint32_t n)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n)
		p[i] = 0.0f;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t value)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(i < n)
		p[i] = value;
}

// =============================================================================================
// Launch wrappers (called from the host runtime)
// =============================================================================================
namespace swk
{
LaunchEvents& launch_events()
{
	static thread_local LaunchEvents e;
	return e;
}
} // namespace swk

// A kernel launch that, when a timed stage offers events (profiling), records them at the kernel's
// own start and end (hipExtLaunchKernelGGL), so the stage's time is the kernel's duration.
template <typename F, typename... Args>
static void timed_launch(F kernel, dim3 grid, dim3 block, uint32_t shm, hipStream_t st, Args... args)
{
	swk::LaunchEvents& e = swk::launch_events();
	if(e.start && !e.used)
	{
		e.used = true;
		hipExtLaunchKernelGGL(kernel, grid, block, shm, st, e.start, e.stop, 0u, args...);
	}
	else
		hipLaunchKernelGGL(kernel, grid, block, shm, st, args...);
}

namespace swk
{

void launch_setup(hipStream_t st, uint32_t lanes, SetupArgs a, const ViewsArg& views, const sw_surface_desc* surfaces,
	uint32_t ns, SurfExt* surf_ext, uint32_t* blk_surf, const MapPtrs& m)
{
	const uint32_t groups = lanes == 256 ? 1u : 4u;
	a.setup_blocks = (a.num_draws + groups - 1) / groups;
	const uint32_t prep_blocks = (ns + 255) / 256;
	const uint32_t blocks = a.setup_blocks + prep_blocks;
	if(blocks == 0)
		return;
	if(lanes == 256)
		timed_launch(k_setup_draws<256>, dim3(blocks), dim3(256), 0, st, a, views, surfaces, ns, surf_ext, blk_surf, m);
	else
		timed_launch(k_setup_draws<64>, dim3(blocks), dim3(256), 0, st, a, views, surfaces, ns, surf_ext, blk_surf, m);
}

template <int LANES>
static void launch_fused_lanes(hipStream_t st, uint32_t color_bits, uint32_t blocks, const SetupArgs& a, const ViewsArg& views,
	const SurfaceArgs& sa, const MapPtrs& m)
{
	if(color_bits == 64)
		timed_launch(k_setup_and_surfaces<LANES, Color64T>, dim3(blocks), dim3(256), 0, st, a, views, sa, m);
	else
		timed_launch(k_setup_and_surfaces<LANES, Color32T>, dim3(blocks), dim3(256), 0, st, a, views, sa, m);
}

void launch_setup_and_surfaces(hipStream_t st, uint32_t lanes, uint32_t color_bits, SetupArgs a, const ViewsArg& views,
	const SurfaceArgs& sa, const MapPtrs& m)
{
	const uint32_t groups = lanes == 256 ? 1u : 4u;
	a.setup_blocks = (a.num_draws + groups - 1) / groups;
	const uint64_t bt = surf_block_texels(sa.arena_texels);
	const uint64_t surf_blocks = sa.num_surfaces ? (sa.arena_texels + bt - 1) / bt : 0;
	const uint64_t blocks = a.setup_blocks + surf_blocks;
	if(blocks == 0)
		return;
	if(lanes == 256)
		launch_fused_lanes<256>(st, color_bits, uint32_t(blocks), a, views, sa, m);
	else
		launch_fused_lanes<64>(st, color_bits, uint32_t(blocks), a, views, sa, m);
}

void launch_build_surfaces(hipStream_t st, uint32_t color_bits, const MapPtrs& m, const SurfaceArgs& sa, const ViewsArg& views,
	const uint32_t* counters)
{
	if(sa.arena_texels == 0 || sa.num_surfaces == 0)
		return;
	const uint64_t bt = surf_block_texels(sa.arena_texels);
	const uint64_t blocks = (sa.arena_texels + bt - 1) / bt;
	const bool sh = sa.num_shadows != 0;
	// blocks of four or more texels per thread stage their lightmap lines in LDS: a line serves the texels
	// of a lightmap cell's width in its row, which then repays the staging's extra step and
	// barrier (with two texels per thread, 1080p shadowed frames, it does not: 9.1 us against 8.9)
	const bool lines = bt >= SURF_LINES_MIN_TEXELS;
	const dim3 g{uint32_t(blocks)}, b{SURF_BLOCK};
	if(color_bits == 64)
	{
		if(sh)
			lines ? timed_launch(k_build_surfaces<Color64T, true, true>, g, b, 0, st, m, sa, views, counters)
				  : timed_launch(k_build_surfaces<Color64T, true, false>, g, b, 0, st, m, sa, views, counters);
		else
			lines ? timed_launch(k_build_surfaces<Color64T, false, true>, g, b, 0, st, m, sa, views, counters)
				  : timed_launch(k_build_surfaces<Color64T, false, false>, g, b, 0, st, m, sa, views, counters);
	}
	else
	{
		if(sh)
			lines ? timed_launch(k_build_surfaces<Color32T, true, true>, g, b, 0, st, m, sa, views, counters)
				  : timed_launch(k_build_surfaces<Color32T, true, false>, g, b, 0, st, m, sa, views, counters);
		else
			lines ? timed_launch(k_build_surfaces<Color32T, false, true>, g, b, 0, st, m, sa, views, counters)
				  : timed_launch(k_build_surfaces<Color32T, false, false>, g, b, 0, st, m, sa, views, counters);
	}
}

template <class CT, bool SSPAN>
static void launch_raster_ct(hipStream_t st, uint32_t tile_rows, const MapPtrs& m, const DrawBounds* bounds, const DrawRec* drec,
	const PartRec* parts, const SpanRec* spans, const ViewsArg& views, const void* arena, const void* skybox,
	void* fb, uint32_t total_tiles, const uint32_t* counters)
{
	typedef typename CT::T C;
	const uint32_t blocks = (total_tiles + 3) / 4;
	if(blocks == 0)
		return;
	const C* a = static_cast<const C*>(arena);
	const C* sky = static_cast<const C*>(skybox);
	C* out = static_cast<C*>(fb);
	if(tile_rows == 4)
		timed_launch(k_raster_tiles<CT, 4, SSPAN>, dim3(blocks), dim3(256), 0, st, m, bounds, drec, parts, spans, views,
			a, sky, out, total_tiles, counters);
	else if(tile_rows == 8)
		timed_launch(k_raster_tiles<CT, 8, SSPAN>, dim3(blocks), dim3(256), 0, st, m, bounds, drec, parts, spans, views,
			a, sky, out, total_tiles, counters);
	else
		timed_launch(k_raster_tiles<CT, 16, SSPAN>, dim3(blocks), dim3(256), 0, st, m, bounds, drec, parts, spans, views,
			a, sky, out, total_tiles, counters);
}

void launch_raster_tiles(hipStream_t st, uint32_t color_bits, uint32_t tile_rows, bool scalar_spans, const MapPtrs& m,
	const DrawBounds* bounds, const DrawRec* drec, const PartRec* parts, const SpanRec* spans, const ViewsArg& views,
	const void* arena, const void* skybox, void* fb, uint32_t total_tiles, const uint32_t* counters)
{
	if(color_bits == 64)
	{
		if(scalar_spans)
			launch_raster_ct<Color64T, true>(st, tile_rows, m, bounds, drec, parts, spans, views, arena, skybox, fb, total_tiles,
				counters);
		else
			launch_raster_ct<Color64T, false>(st, tile_rows, m, bounds, drec, parts, spans, views, arena, skybox, fb, total_tiles,
				counters);
	}
	else if(scalar_spans)
		launch_raster_ct<Color32T, true>(st, tile_rows, m, bounds, drec, parts, spans, views, arena, skybox, fb, total_tiles,
			counters);
	else
		launch_raster_ct<Color32T, false>(st, tile_rows, m, bounds, drec, parts, spans, views, arena, skybox, fb, total_tiles,
			counters);
}

void launch_finish(hipStream_t st, const DrawRec* drec, uint32_t nd, uint32_t* counters, uint32_t* result)
{
	hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, st, drec, nd, counters, result);
}

void launch_bloom(hipStream_t st, const BloomArgs& a, bool average)
{
	switch(a.scale_log2)
	{
	case 1: hipLaunchKernelGGL(k_bloom_downscale<2>, dim3(a.bh), dim3(256), 0, st, a); break;
	case 2: hipLaunchKernelGGL(k_bloom_downscale<4>, dim3(a.bh), dim3(256), 0, st, a); break;
	default: hipLaunchKernelGGL(k_bloom_downscale<8>, dim3(a.bh), dim3(256), 0, st, a); break;
	}
	if(a.radius > 0)
	{
		const dim3 blur_grid((a.bw + 256 * BLUR_PX - 1) / (256 * BLUR_PX), a.bh);
		hipLaunchKernelGGL(k_bloom_blur, blur_grid, dim3(256), 0, st, a, 0u, average ? 1u : 0u);
		hipLaunchKernelGGL(k_bloom_blur, blur_grid, dim3(256), 0, st, a, 1u, 0u);
	}
	else if(average)
		hipLaunchKernelGGL(k_bloom_average, dim3(1), dim3(256), 0, st, a);
	const dim3 grid((a.w + 256 * TM_PX - 1) / (256 * TM_PX), a.h);
	if(a.linear)
		hipLaunchKernelGGL(k_bloom_tonemap<true>, grid, dim3(256), 0, st, a);
	else
		hipLaunchKernelGGL(k_bloom_tonemap<false>, grid, dim3(256), 0, st, a);
}

void launch_portal_mix(hipStream_t st, uint32_t color_bits, const uint32_t* lite, const PortalMixRec* recs, uint32_t num,
	uint64_t max_texels, void* texels)
{
	if(num == 0 || max_texels == 0)
		return;
	const dim3 grid(uint32_t((max_texels + PORTAL_MIX_BLOCK_TEXELS - 1) / PORTAL_MIX_BLOCK_TEXELS), num);
	if(color_bits == 64)
		hipLaunchKernelGGL(k_portal_mix<Color64T>, grid, dim3(256), 0, st, lite, recs, static_cast<uint64_t*>(texels));
	else
		hipLaunchKernelGGL(k_portal_mix<Color32T>, grid, dim3(256), 0, st, lite, recs, static_cast<uint32_t*>(texels));
}

void launch_tonemap(hipStream_t st, const uint64_t* hdr, uint32_t w, uint32_t h, float isr, float isg, float isb,
	uint32_t* dst, uint32_t dst_pitch, float* partials, float* totals)
{
	const uint32_t blocks = std::max(1u, std::min(h, TONEMAP_BLOCKS));
	hipLaunchKernelGGL(k_tonemap, dim3(blocks), dim3(256), 0, st, hdr, w, h, isr, isg, isb, dst, dst_pitch, partials);
	if(totals)
		hipLaunchKernelGGL(k_tonemap_sums, dim3(1), dim3(256), 0, st, partials, blocks, totals);
}

} // namespace swk
