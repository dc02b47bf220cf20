// CDNA4 kernels of the surface builder, span rasteriser and tonemap.
//
// Pipeline for a batch of views (the plain path is a batch of one view):
//   k_prep_surfaces   one thread per surface: global arena start/end, ordering check
//   k_build_surfaces  one thread per arena texel (flattened over all surfaces of all views):
//                     surfaces.rs:433-852 + :854-971 for that texel
//   k_setup_draws     one wave per draw: the edge walk of rasterizer.rs:137-271 (or the triangle
//                     split of :765-982) executed wave-uniformly; each part's rows are spread over
//                     the lanes, which write one 32-byte span record per (part, row) holding the
//                     span's clipped x range and the start/step of its interpolants
//                     (:316-422, :500-556, :673-703, :1000-1214)
//   k_bin_bands       one workgroup per (view, band of BAND_H rows): ordered compaction of the
//                     draws overlapping the band (ballot + popcount prefix per wave, LDS across waves)
//   k_raster_tiles    one wave per 64 x BAND_H screen tile: tile colour in LDS, walks the band's
//                     draw list in draw order, per pixel: texel address from the span record, fetch,
//                     write_into_framebuffer (:424-445, :558-572, :705-719, :1223-1279)
//   k_tonemap         one thread per pixel: postprocessor.rs:237-253
#include "sw_device.h"
#include "sw_internal.h"

using namespace swd;

// =============================================================================================
// Surfaces
// =============================================================================================

__global__ void k_prep_surfaces(
	const sw_surface_desc* __restrict__ surfaces, uint32_t num_surfaces, const ViewExt* __restrict__ views,
	uint32_t num_views, uint64_t* __restrict__ surf_end, uint64_t arena_texels, uint32_t num_levels,
	const sw_texture_level* __restrict__ levels, uint64_t m_num_lm_simple, uint64_t m_num_lm_dir,
	uint32_t* __restrict__ counters)
{
	const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if(s == 0)
	{
		// reset per-batch counters (no memset nodes: this runs inside captured graphs too)
		for(int i = 0; i < COUNTER_COUNT; ++i)
			counters[i] = 0;
	}
	if(s >= num_surfaces)
		return;
	// view owning surface s (views are sorted by first_surface)
	uint32_t lo = 0, hi = num_views;
	while(hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if(views[mid].v.first_surface <= s)
			lo = mid;
		else
			hi = mid;
	}
	const sw_surface_desc& d = surfaces[s];
	const uint64_t start = views[lo].v.surface_arena_offset + d.pixels_offset;
	const uint64_t end = start + uint64_t(d.size[0]) * d.size[1];
	surf_end[s] = end;
	bool bad = end > arena_texels || d.size[0] > 65536u || d.size[1] > 65536u;
	if((d.flags & SW_SURF_DIFFUSE_LAYER) && (d.texture_level >= num_levels || (levels[d.texture_level].flags & SW_TEX_LITE_LEVEL)))
		bad = true;
	if((d.flags & SW_SURF_EMISSIVE) && d.emissive_level != SW_NO_TEXTURE &&
		(d.emissive_level >= num_levels || !(levels[d.emissive_level].flags & SW_TEX_LITE_LEVEL)))
		bad = true;
	if((d.flags & SW_SURF_DIFFUSE_LAYER) && d.lightmap_offset != 0 && d.size[0] > 0 && d.size[1] > 0)
	{
		// Every lightmap fetch (rows lv, lv + 1; columns lu, lu + 1) must stay inside the polygon's
		// lightmap (the reference's debug_assert!s at surfaces.rs:482, :541) and inside the buffer.
		const uint32_t sl = d.lightmap_scale_log2;
		const uint64_t num_lm = (d.flags & SW_SURF_DIRECTIONAL) ? m_num_lm_dir : m_num_lm_simple;
		if(sl > 4 || d.lightmap_size[0] < 2 || d.lightmap_size[1] < 2 ||
			((uint64_t(d.size[0]) - 1 + d.lightmap_tc_shift[0]) >> sl) + 1 >= d.lightmap_size[0] ||
			((uint64_t(d.size[1]) - 1 + d.lightmap_tc_shift[1]) >> sl) + 1 >= d.lightmap_size[1] ||
			uint64_t(d.lightmap_offset) + uint64_t(d.lightmap_size[0]) * d.lightmap_size[1] > num_lm)
			bad = true;
	}
	if(s > 0)
	{
		// Arena ranges must follow descriptor order (the reference's bump allocator,
		// partial_renderer.rs:1617-1618). Recompute the predecessor's end to check.
		uint32_t plo = 0, phi = num_views;
		while(phi - plo > 1)
		{
			const uint32_t mid = (plo + phi) >> 1;
			if(views[mid].v.first_surface <= s - 1)
				plo = mid;
			else
				phi = mid;
		}
		const sw_surface_desc& p = surfaces[s - 1];
		const uint64_t pend = views[plo].v.surface_arena_offset + p.pixels_offset + uint64_t(p.size[0]) * p.size[1];
		if(pend > start)
			bad = true;
	}
	if(bad)
		atomicOr(&counters[COUNTER_ERROR], ERROR_BAD_SURFACE);
}

struct LmE
{
	float amb[3];
	V3 dir;
	float dev;
	float col[3];
};

SW_DEV LmE lm_load(const MapPtrs& m, bool directional, uint64_t i)
{
	LmE e;
	if(directional)
	{
		const sw_directional_lightmap_element& d = m.lm_dir[i];
		e.amb[0] = d.ambient_light[0];
		e.amb[1] = d.ambient_light[1];
		e.amb[2] = d.ambient_light[2];
		e.dir = v3(d.light_direction_vector_scaled[0], d.light_direction_vector_scaled[1], d.light_direction_vector_scaled[2]);
		e.dev = d.directional_light_deviation;
		e.col[0] = d.directional_light_color[0];
		e.col[1] = d.directional_light_color[1];
		e.col[2] = d.directional_light_color[2];
	}
	else
	{
		e.amb[0] = m.lm_simple[i * 3 + 0];
		e.amb[1] = m.lm_simple[i * 3 + 1];
		e.amb[2] = m.lm_simple[i * 3 + 2];
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

// One texel of build_surface_impl_6_static_params (surfaces.rs:433-852), specialisations as
// runtime flags that are uniform per surface.
template <class CT>
SW_DEV typename CT::T surface_texel(
	const MapPtrs& m, const sw_surface_desc& s, const sw_light* __restrict__ lights,
	const uint32_t* __restrict__ light_indices, uint32_t num_light_indices, uint32_t num_lights_total, V3 cam_pos,
	uint32_t dst_u, uint32_t dst_v)
{
	const sw_texture_level tex = m.levels[s.texture_level];
	const bool use_normal_map = (tex.flags & SW_TEX_HAS_NORMAL_MAP) != 0;
	const uint32_t spec = (tex.flags & SW_TEX_HAS_NON_ONE_ROUGHNESS) ? ((tex.flags & SW_TEX_IS_METAL) ? 2u : 1u) : 0u;
	const bool directional = (s.flags & SW_SURF_DIRECTIONAL) != 0;
	const bool has_lightmap = s.lightmap_offset != 0;
	const uint32_t num_lights = s.num_lights < SW_MAX_POLYGON_LIGHTS ? s.num_lights : SW_MAX_POLYGON_LIGHTS;

	const V3 u_vec = v3(s.basis_u[0], s.basis_u[1], s.basis_u[2]);
	const V3 v_vec = v3(s.basis_v[0], s.basis_v[1], s.basis_v[2]);
	const V3 basis_normal = v3(s.basis_normal[0], s.basis_normal[1], s.basis_normal[2]);
	// :456-457
	const V3 start_pos = add(add(v3(s.basis_start[0], s.basis_start[1], s.basis_start[2]),
								 scale(u_vec, float(s.tc_start[0]) + 0.5f)),
		scale(v_vec, float(s.tc_start[1]) + 0.5f));
	const V3 u_n = scale(u_vec, inv_sqrt_fast(fmax_r(len2(u_vec), MIN_POSITIVE_VALUE)));
	const V3 v_n = scale(v_vec, inv_sqrt_fast(fmax_r(len2(v_vec), MIN_POSITIVE_VALUE)));

	// texel fetch with rem_euclid wrap (:507-517, :843-847)
	const int32_t tw = int32_t(tex.width), th = int32_t(tex.height);
	const int32_t src_v = rem_euclid(s.tc_start[1] + int32_t(dst_v), th);
	const int32_t src_u = rem_euclid(rem_euclid(s.tc_start[0], tw) + int32_t(dst_u % uint32_t(tw)), tw);
	const sw_texel texel = m.texels[tex.offset + uint64_t(uint32_t(src_v)) * uint32_t(tw) + uint32_t(src_u)];

	const V3 start_pos_v = scalar_mul_add(v_vec, float(dst_v), start_pos);
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
		const uint32_t sl = s.lightmap_scale_log2;
		const float inv_scale = 1.0f / float(1u << sl);
		const float k_shift = 0.5f * inv_scale;
		const uint32_t mask = (1u << sl) - 1u;
		// vertical mix of the two lightmap rows (:477-502) for the two columns this texel needs
		const uint32_t base_v = dst_v + s.lightmap_tc_shift[1];
		const uint32_t lv = base_v >> sl;
		const float kv = mul_add(float(base_v & mask), inv_scale, k_shift);
		const uint32_t base_u = dst_u + s.lightmap_tc_shift[0];
		const uint32_t lu = base_u >> sl;
		const float ku = mul_add(float(base_u & mask), inv_scale, k_shift);
		const uint64_t row0 = uint64_t(s.lightmap_offset) + uint64_t(lv) * s.lightmap_size[0];
		const uint64_t row1 = row0 + s.lightmap_size[0];
		const LmE l00 = lm_load(m, directional, row0 + lu);
		const LmE l01 = lm_load(m, directional, row0 + lu + 1);
		const LmE l10 = lm_load(m, directional, row1 + lu);
		const LmE l11 = lm_load(m, directional, row1 + lu + 1);
		const LmE line0 = lm_mix(l10, l00, kv, directional);
		const LmE line1 = lm_mix(l11, l01, kv, directional);
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
			const V3 to_cam = sub(cam_pos, pos);
			const V3 to_cam_ts = v3(dot(to_cam, u_n), dot(to_cam, v_n), dot(to_cam, basis_normal));
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
			const V3 to_cam = sub(cam_pos, pos);
			const float n_dot = dot(to_cam, normal);
			refl = sub(scale(normal, 2.0f * n_dot), to_cam);
			refl_len2 = fmax_r(len2(refl), MIN_POSITIVE_VALUE);
			const float cos_a = fmax_r(n_dot * inv_sqrt_fast(refl_len2), 0.0f);
			sk = specular_k(spec, fresnel_factor_base(cos_a), tr);
			inv_r = inv_fast(tr);
		}
		for(uint32_t li = 0; li < num_lights; ++li)
		{
			const uint32_t idx_pos = s.first_light + li;
			if(idx_pos >= num_light_indices)
				break;
			const uint32_t lidx = light_indices[idx_pos];
			if(lidx >= num_lights_total)
				continue;
			const sw_light L = lights[lidx];
			const V3 to_light = sub(v3(L.position[0], L.position[1], L.position[2]), pos);
			const float tl2 = fmax_r(len2(to_light), MIN_POSITIVE_VALUE);
			const float dist_f = fmax_r(inv_fast(tl2) - L.inv_square_radius, 0.0f);
			const float sdf = 1.0f * dist_f; // shadow factor 1: ShadowMap::None
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

// surfaces.rs:903-971, one texel
template <class CT>
SW_DEV typename CT::T emissive_texel(const MapPtrs& m, const sw_surface_desc& s, uint32_t blend, typename CT::T dst,
	uint32_t dst_u, uint32_t dst_v)
{
	const sw_texture_level tex = m.levels[s.emissive_level];
	const int32_t tw = int32_t(tex.width), th = int32_t(tex.height);
	const int32_t src_v = rem_euclid(s.tc_start[1] + int32_t(dst_v), th);
	const int32_t src_u = rem_euclid(rem_euclid(s.tc_start[0], tw) + int32_t(dst_u % uint32_t(tw)), tw);
	const uint32_t t = m.lite[tex.offset + uint64_t(uint32_t(src_v)) * uint32_t(tw) + uint32_t(src_u)];
	// light_vec = cvtps2dq(light * 256): lanes [B, G, R]
	const int32_t lb = f32_to_i32_rne_x86(s.emissive_light[2] * 256.0f);
	const int32_t lg = f32_to_i32_rne_x86(s.emissive_light[1] * 256.0f);
	const int32_t lr = f32_to_i32_rne_x86(s.emissive_light[0] * 256.0f);
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

template <class CT>
__global__ __launch_bounds__(256) void k_build_surfaces(
	MapPtrs m, const sw_surface_desc* __restrict__ surfaces, uint32_t num_surfaces, const uint64_t* __restrict__ surf_end,
	const ViewExt* __restrict__ views, uint32_t num_views, const sw_light* __restrict__ lights, uint32_t num_lights,
	const uint32_t* __restrict__ light_indices, uint32_t num_light_indices, typename CT::T* __restrict__ arena,
	uint64_t arena_texels, const uint32_t* __restrict__ counters)
{
	if(counters[COUNTER_ERROR] & ERROR_BAD_SURFACE)
		return;
	const uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	// Wave-level search for the first surface whose end lies past the wave's first texel.
	const uint64_t wave_first = idx & ~uint64_t(63);
	uint32_t s0 = 0;
	if((threadIdx.x & 63) == 0)
	{
		uint32_t lo = 0, hi = num_surfaces;
		while(lo < hi)
		{
			const uint32_t mid = (lo + hi) >> 1;
			if(surf_end[mid] <= wave_first)
				lo = mid + 1;
			else
				hi = mid;
		}
		s0 = lo;
	}
	s0 = __shfl(s0, 0);
	if(idx >= arena_texels)
		return;
	uint32_t s = s0;
	while(s < num_surfaces && surf_end[s] <= idx)
		++s;
	if(s >= num_surfaces)
		return;
	const sw_surface_desc& d = surfaces[s];
	const uint64_t size = uint64_t(d.size[0]) * d.size[1];
	const uint64_t start = surf_end[s] - size;
	if(idx < start)
		return; // gap between surfaces
	uint32_t lo = 0, hi = num_views;
	while(hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if(views[mid].v.first_surface <= s)
			lo = mid;
		else
			hi = mid;
	}
	const V3 cam = v3(views[lo].v.cam_pos[0], views[lo].v.cam_pos[1], views[lo].v.cam_pos[2]);
	const uint32_t local = uint32_t(idx - start);
	const uint32_t du = local % d.size[0];
	const uint32_t dv = local / d.size[0];
	typename CT::T out = arena[idx];
	if(d.flags & SW_SURF_DIFFUSE_LAYER)
		out = surface_texel<CT>(m, d, lights, light_indices, num_light_indices, num_lights, cam, du, dv);
	if((d.flags & SW_SURF_EMISSIVE) && d.emissive_level < m.num_levels)
		out = emissive_texel<CT>(m, d, d.emissive_blend, out, du, dv);
	arena[idx] = out;
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

struct SetupState
{
	uint32_t draw;
	uint32_t num_parts;
	int32_t row_min, row_max, x_min, x_max;
	uint64_t pixels, blend_pixels;
};

// Emits one part: rows [y_start_int, y_end_int) spread over the wave's lanes.
// kind: 0 perspective, 1 line z, 2 affine, 3 triangle.
SW_DEV void emit_part(const sw_draw& d, const PolyEq& e, uint32_t kind, const int32_t clip[4], int32_t y_start,
	int32_t y_end, Side ls, Side rs, const AffinePart* ap, const TriPart* tp, PartRec* __restrict__ parts,
	SpanRec* __restrict__ spans, uint32_t* __restrict__ counters, uint64_t span_capacity, uint32_t part_slot_base,
	SetupState& st)
{
	const int32_t y_start_int = imax32(fixed16_round_to_int(y_start), clip[1]);
	const int32_t y_end_int = imin32(fixed16_round_to_int(y_end), clip[3]);
	if(y_start_int >= y_end_int)
		return; // no rows: the reference draws nothing for this part
	if(st.num_parts >= MAX_PARTS_PER_DRAW)
		return;
	const int32_t y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y_start;
	const int32_t x_left0 = ls.x_start + fixed16_mul(y_start_delta, ls.dx_dy) + FIXED16_HALF;
	const int32_t x_right0 = rs.x_start + fixed16_mul(y_start_delta, rs.dx_dy) + FIXED16_HALF;
	const uint32_t rows = uint32_t(y_end_int - y_start_int);
	const uint32_t lane = threadIdx.x & 63;

	uint64_t base = 0;
	if(lane == 0)
		base = atomicAdd(reinterpret_cast<unsigned long long*>(&counters[COUNTER_SPANS_LO]), (unsigned long long)rows);
	base = __shfl(base, 0);
	const bool fits = base + rows <= span_capacity;

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
	if(lane == 0)
		parts[part_slot_base + st.num_parts] = pr;
	st.num_parts++;
	if(!fits)
	{
		if(lane == 0)
			atomicOr(&counters[COUNTER_ERROR], ERROR_SPAN_OVERFLOW);
		return;
	}

	int32_t lx_min = INT32_MAX, lx_max = INT32_MIN, ly_min = INT32_MAX, ly_max = INT32_MIN;
	uint64_t lpix = 0;
	for(uint32_t k = lane; k < rows; k += 64)
	{
		const int32_t y = y_start_int + int32_t(k);
		const int32_t x_left = wadd(x_left0, wmul(int32_t(k), ls.dx_dy));
		const int32_t x_right = wadd(x_right0, wmul(int32_t(k), rs.dx_dy));
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
			lx_min = imin32(lx_min, x_start_int);
			lx_max = imax32(lx_max, x_end_int);
			ly_min = imin32(ly_min, y);
			ly_max = imax32(ly_max, y + 1);
			lpix += uint64_t(x_end_int - x_start_int);
		}
		spans[base + k] = r;
	}
	// wave reductions
	for(int off = 32; off > 0; off >>= 1)
	{
		lx_min = imin32(lx_min, __shfl_xor(lx_min, off));
		lx_max = imax32(lx_max, __shfl_xor(lx_max, off));
		ly_min = imin32(ly_min, __shfl_xor(ly_min, off));
		ly_max = imax32(ly_max, __shfl_xor(ly_max, off));
		lpix += __shfl_xor(lpix, off);
	}
	st.x_min = imin32(st.x_min, lx_min);
	st.x_max = imax32(st.x_max, lx_max);
	st.row_min = imin32(st.row_min, ly_min);
	st.row_max = imax32(st.row_max, ly_max);
	st.pixels += lpix;
}

__global__ __launch_bounds__(256) void k_setup_draws(const sw_draw* __restrict__ draws, uint32_t num_draws,
	const int32_t* __restrict__ vdata, uint64_t num_vertex_ints, const ViewExt* __restrict__ views, uint32_t num_views,
	PartRec* __restrict__ parts, SpanRec* __restrict__ spans, DrawRec* __restrict__ drec, uint32_t* __restrict__ counters,
	uint64_t span_capacity, uint64_t arena_texels, uint64_t skybox_texels, uint64_t lite_texels)
{
	const uint32_t di = blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
	if(di >= num_draws)
		return;
	const uint32_t lane = threadIdx.x & 63;
	const sw_draw d = draws[di];
	uint32_t lo = 0, hi = num_views;
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
	const int32_t clip[4] = {imax32(d.clip_rect[0], 0), imax32(d.clip_rect[1], 0), imin32(d.clip_rect[2], int32_t(view.width)),
		imin32(d.clip_rect[3], int32_t(view.height))};

	SetupState st;
	st.draw = di;
	st.num_parts = 0;
	st.row_min = INT32_MAX;
	st.row_max = INT32_MIN;
	st.x_min = INT32_MAX;
	st.x_max = INT32_MIN;
	st.pixels = 0;
	st.blend_pixels = 0;
	const uint32_t part_slot_base = di * MAX_PARTS_PER_DRAW;

	// Texture and vertex ranges must lie inside their buffers (the reference indexes them unchecked
	// in release builds, fast_math.rs:750-777; here an invalid draw is skipped instead).
	const uint64_t tex_texels = uint64_t(uint32_t(d.texture_size[0])) * uint32_t(d.texture_size[1]);
	bool valid = d.texture_size[0] > 0 && d.texture_size[1] > 0 && d.texture_size[0] <= 65536 && d.texture_size[1] <= 65536;
	if(d.kind == SW_DRAW_POLYGON)
	{
		const uint64_t limit = d.texture_source == SW_TEXSRC_SKYBOX ? skybox_texels : arena_texels;
		const uint64_t base = d.texture_source == SW_TEXSRC_SKYBOX ? 0 : view.surface_arena_offset;
		valid = valid && d.num_vertices >= 3 && d.num_vertices <= SW_MAX_POLYGON_VERTICES &&
			uint64_t(d.first_vertex_int) + 2ull * d.num_vertices <= num_vertex_ints &&
			base + d.texture_offset + tex_texels <= limit && d.blend <= SW_BLEND_ALPHA_BLEND;
	}
	else
		valid = valid && d.kind == SW_DRAW_TRIANGLE && uint64_t(d.first_vertex_int) + 21ull <= num_vertex_ints &&
			d.texture_offset + tex_texels <= lite_texels && d.blend <= SW_BLEND_ALPHA_BLEND;

	if(valid && d.kind == SW_DRAW_POLYGON)
	{
		// rasterizer.rs:137-271, executed identically by every lane
		const uint32_t n = d.num_vertices;
		const int32_t* v = vdata + d.first_vertex_int;
		const PolyEq e = make_poly_eq(d);
		const uint32_t kind = d.mode == SW_TC_LINE_Z ? 1u : (d.mode == SW_TC_AFFINE ? 2u : 0u);
		uint32_t lower = 0;
		int32_t min_y = v[1];
		for(uint32_t i = 0; i < n; ++i)
			if(v[i * 2 + 1] < min_y)
			{
				min_y = v[i * 2 + 1];
				lower = i;
			}
		uint32_t left = lower, right = lower;
		int32_t cur_y = min_y;
		for(uint32_t iter = 0; iter < n; ++iter)
		{
			uint32_t nl = left + n - 1;
			if(nl >= n)
				nl -= n;
			uint32_t nr = right + 1;
			if(nr >= n)
				nr -= n;
			const int32_t lx = v[left * 2], ly = v[left * 2 + 1], nlx = v[nl * 2], nly = v[nl * 2 + 1];
			const int32_t rx = v[right * 2], ry = v[right * 2 + 1], nrx = v[nr * 2], nry = v[nr * 2 + 1];
			const int32_t dy_left = nly - ly;
			const int32_t dy_right = nry - ry;
			const int32_t next_y = imin32(nly, nry);
			bool have = false;
			Side ls, rs;
			if(dy_left > FIXED16_HALF && dy_right > FIXED16_HALF)
			{
				const int32_t dxl = fixed16_div(nlx - lx, dy_left);
				const int32_t dxr = fixed16_div(nrx - rx, dy_right);
				ls = Side{lx + fixed16_mul(dxl, cur_y - ly), dxl};
				rs = Side{rx + fixed16_mul(dxr, cur_y - ry), dxr};
				have = true;
			}
			else if(dy_left > 0 && dy_right > 0)
			{
				const int32_t cy = fixed16_round_to_int(cur_y);
				const int32_t ny = fixed16_round_to_int(next_y);
				if(cy < ny)
				{
					const int32_t thin = int_to_fixed16(cy) + FIXED16_HALF;
					ls = Side{lx + fixed16_mul_div(thin - ly, nlx - lx, dy_left), 0};
					rs = Side{rx + fixed16_mul_div(thin - ry, nrx - rx, dy_right), 0};
					have = true;
				}
			}
			if(have)
			{
				if(kind == 2)
				{
					const int32_t y_start_int = imax32(fixed16_round_to_int(cur_y), clip[1]);
					const int32_t y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - cur_y;
					const AffinePart ap = affine_part_setup(d, cur_y, next_y, ls, rs, y_start_delta);
					emit_part(d, e, kind, clip, cur_y, next_y, ls, rs, &ap, nullptr, parts, spans, counters, span_capacity,
						part_slot_base, st);
				}
				else
					emit_part(d, e, kind, clip, cur_y, next_y, ls, rs, nullptr, nullptr, parts, spans, counters,
						span_capacity, part_slot_base, st);
			}
			if(nl == nr)
				break;
			if(nry < nly)
				right = nr;
			else
				left = nl;
			cur_y = next_y;
		}
	}
	else if(valid)
	{
		// rasterizer.rs:765-982
		const int32_t* p = vdata + d.first_vertex_int;
		int32_t vx[3], vy[3], vtc[3][2], vl[3][3];
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
		const int32_t long_edge_dy = imax32(vy[up] - vy[lw], FIXED16_HALF);
		const int32_t lower_part_dy = imax32(vy[mi] - vy[lw], FIXED16_HALF);
		const int32_t upper_part_dy = imax32(vy[up] - vy[mi], FIXED16_HALF);
		const int32_t long_edge_dx_dy = fixed16_div(vx[up] - vx[lw], long_edge_dy);
		const int32_t long_edge_x_in_middle = vx[lw] + fixed16_mul(long_edge_dx_dy, vy[mi] - vy[lw]);
		const int32_t lower_part_dx_dy = fixed16_div(vx[mi] - vx[lw], lower_part_dy);
		const int32_t upper_part_dx_dy = fixed16_div(vx[up] - vx[mi], upper_part_dy);
		int32_t le_dtc[2], le_tc_mid[2], dtc_lower[2], dtc_upper[2];
		for(int i = 0; i < 2; ++i)
		{
			le_dtc[i] = fixed16_div(vtc[up][i] - vtc[lw][i], long_edge_dy);
			le_tc_mid[i] = vtc[lw][i] + fixed16_mul(le_dtc[i], lower_part_dy);
			dtc_lower[i] = fixed16_div(vtc[mi][i] - vtc[lw][i], lower_part_dy);
			dtc_upper[i] = fixed16_div(vtc[up][i] - vtc[mi][i], upper_part_dy);
		}
		int32_t le_dl[3], le_l_mid[3], dl_lower[3], dl_upper[3];
		for(int i = 0; i < 3; ++i)
		{
			le_dl[i] = fixed16_div(vl[up][i] - vl[lw][i], long_edge_dy);
			le_l_mid[i] = vl[lw][i] + fixed16_mul(le_dl[i], lower_part_dy);
			dl_lower[i] = fixed16_div(vl[mi][i] - vl[lw][i], lower_part_dy);
			dl_upper[i] = fixed16_div(vl[up][i] - vl[mi][i], upper_part_dy);
		}
		const PolyEq e = PolyEq{};
		int32_t y0[2], y1[2], tcs[2][2], dtcy[2][2], ls_[3][2], dly[2][3];
		Side lsd[2], rsd[2];
		int32_t d_tc_dx[2], d_light_dx[3];
		if(long_edge_x_in_middle >= vx[mi])
		{
			const int32_t middle_dx = imax32(long_edge_x_in_middle - vx[mi], FIXED16_HALF);
			for(int i = 0; i < 2; ++i)
				d_tc_dx[i] = fixed16_div(le_tc_mid[i] - vtc[mi][i], middle_dx);
			for(int i = 0; i < 3; ++i)
				d_light_dx[i] = fixed16_div(le_l_mid[i] - vl[mi][i], middle_dx);
			y0[0] = vy[lw];
			y1[0] = vy[mi];
			lsd[0] = Side{vx[lw], lower_part_dx_dy};
			rsd[0] = Side{vx[lw], long_edge_dx_dy};
			y0[1] = vy[mi];
			y1[1] = vy[up];
			lsd[1] = Side{vx[mi], upper_part_dx_dy};
			rsd[1] = Side{long_edge_x_in_middle, long_edge_dx_dy};
			for(int i = 0; i < 2; ++i)
			{
				tcs[0][i] = vtc[lw][i];
				dtcy[0][i] = dtc_lower[i];
				tcs[1][i] = vtc[mi][i];
				dtcy[1][i] = dtc_upper[i];
			}
			for(int i = 0; i < 3; ++i)
			{
				ls_[i][0] = vl[lw][i];
				dly[0][i] = dl_lower[i];
				ls_[i][1] = vl[mi][i];
				dly[1][i] = dl_upper[i];
			}
		}
		else
		{
			const int32_t middle_dx = imax32(vx[mi] - long_edge_x_in_middle, FIXED16_HALF);
			for(int i = 0; i < 2; ++i)
				d_tc_dx[i] = fixed16_div(vtc[mi][i] - le_tc_mid[i], middle_dx);
			for(int i = 0; i < 3; ++i)
				d_light_dx[i] = fixed16_div(vl[mi][i] - le_l_mid[i], middle_dx);
			y0[0] = vy[lw];
			y1[0] = vy[mi];
			lsd[0] = Side{vx[lw], long_edge_dx_dy};
			rsd[0] = Side{vx[lw], lower_part_dx_dy};
			y0[1] = vy[mi];
			y1[1] = vy[up];
			lsd[1] = Side{long_edge_x_in_middle, long_edge_dx_dy};
			rsd[1] = Side{vx[mi], upper_part_dx_dy};
			for(int i = 0; i < 2; ++i)
			{
				tcs[0][i] = vtc[lw][i];
				dtcy[0][i] = le_dtc[i];
				tcs[1][i] = le_tc_mid[i];
				dtcy[1][i] = le_dtc[i];
			}
			for(int i = 0; i < 3; ++i)
			{
				ls_[i][0] = vl[lw][i];
				dly[0][i] = le_dl[i];
				ls_[i][1] = le_l_mid[i];
				dly[1][i] = le_dl[i];
			}
		}
		for(int pi = 0; pi < 2; ++pi)
		{
			const int32_t y_start_int = imax32(fixed16_round_to_int(y0[pi]), clip[1]);
			const int32_t y_end_int = imin32(fixed16_round_to_int(y1[pi]), clip[3]);
			if(y_start_int >= y_end_int)
				continue;
			const int32_t y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y0[pi];
			int32_t tc_start[2] = {tcs[pi][0], tcs[pi][1]};
			int32_t dty[2] = {dtcy[pi][0], dtcy[pi][1]};
			int32_t dtx[2] = {d_tc_dx[0], d_tc_dx[1]};
			int32_t l_start[3] = {ls_[0][pi], ls_[1][pi], ls_[2][pi]};
			int32_t dl[3] = {dly[pi][0], dly[pi][1], dly[pi][2]};
			int32_t dlx[3] = {d_light_dx[0], d_light_dx[1], d_light_dx[2]};
			TriPart tp;
			tri_part_setup(d.texture_size, y0[pi], y1[pi], lsd[pi], rsd[pi], y_start_delta, tc_start, dty, dtx, l_start, dl, dlx, tp);
			emit_part(d, e, 3u, clip, y0[pi], y1[pi], lsd[pi], rsd[pi], nullptr, &tp, parts, spans, counters, span_capacity,
				part_slot_base, st);
		}
	}

	if(lane == 0)
	{
		DrawRec r;
		r.num_parts = st.num_parts;
		r.row_min = st.row_min;
		r.row_max = st.row_max;
		r.x_min = st.x_min;
		r.x_max = st.x_max;
		drec[di] = r;
		if(st.pixels)
		{
			const bool reads = d.blend == SW_BLEND_AVERAGE || d.blend == SW_BLEND_ADDITIVE || d.blend == SW_BLEND_ALPHA_TEST ||
				d.blend == SW_BLEND_ALPHA_BLEND;
			atomicAdd(reinterpret_cast<unsigned long long*>(
						  &counters[d.kind == SW_DRAW_POLYGON ? COUNTER_POLY_PIXELS_LO : COUNTER_TRI_PIXELS_LO]),
				(unsigned long long)st.pixels);
			if(reads)
				atomicAdd(reinterpret_cast<unsigned long long*>(&counters[COUNTER_BLEND_PIXELS_LO]), (unsigned long long)st.pixels);
		}
		atomicAdd(&counters[COUNTER_PARTS], st.num_parts);
	}
}

// =============================================================================================
// Band binning: ordered list of draws per (view, band)
// =============================================================================================
__global__ __launch_bounds__(256) void k_bin_bands(const DrawRec* __restrict__ drec, const ViewExt* __restrict__ views,
	uint32_t num_views, uint32_t* __restrict__ band_lists, uint32_t* __restrict__ band_counts,
	const uint32_t* __restrict__ counters)
{
	__shared__ uint32_t wave_counts[4];
	__shared__ uint32_t running;
	const uint32_t band_global = blockIdx.x;
	if(counters[COUNTER_ERROR] & ERROR_SPAN_OVERFLOW)
		return;
	uint32_t lo = 0, hi = num_views;
	while(hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if(views[mid].first_band <= band_global)
			lo = mid;
		else
			hi = mid;
	}
	const ViewExt& ve = views[lo];
	const uint32_t band = band_global - ve.first_band;
	const int32_t y0 = int32_t(band * BAND_H);
	const int32_t y1 = y0 + BAND_H;
	uint32_t* list = band_lists + ve.band_list_base + uint64_t(band) * ve.v.num_draws;
	if(threadIdx.x == 0)
		running = 0;
	__syncthreads();
	const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	for(uint32_t base = 0; base < ve.v.num_draws; base += 256)
	{
		const uint32_t i = base + threadIdx.x;
		bool flag = false;
		if(i < ve.v.num_draws)
		{
			const DrawRec r = drec[ve.v.first_draw + i];
			flag = r.num_parts > 0 && r.row_min < y1 && r.row_max > y0 && r.x_min < r.x_max;
		}
		const uint64_t ballot = __ballot(flag);
		const uint32_t prefix = __popcll(ballot & ((1ull << lane) - 1ull));
		if(lane == 0)
			wave_counts[wave] = __popcll(ballot);
		__syncthreads();
		uint32_t off = running;
		for(uint32_t w = 0; w < wave; ++w)
			off += wave_counts[w];
		if(flag)
			list[off + prefix] = ve.v.first_draw + i;
		__syncthreads();
		if(threadIdx.x == 0)
			running += wave_counts[0] + wave_counts[1] + wave_counts[2] + wave_counts[3];
		__syncthreads();
	}
	if(threadIdx.x == 0)
		band_counts[band_global] = running;
}

// =============================================================================================
// Tiled raster: one wave per 64 x BAND_H tile, colour kept in LDS
// =============================================================================================
template <class CT>
SW_DEV typename CT::T fetch_checked(const typename CT::T* __restrict__ tex, uint32_t u, uint32_t v, const sw_draw& d, bool& ok)
{
	ok = u < uint32_t(d.texture_size[0]) && v < uint32_t(d.texture_size[1]);
	return ok ? tex[u + v * uint32_t(d.texture_size[0])] : typename CT::T(0);
}

template <class CT>
__global__ __launch_bounds__(256) void k_raster_tiles(MapPtrs m, const sw_draw* __restrict__ draws,
	const DrawRec* __restrict__ drec, const PartRec* __restrict__ parts, const SpanRec* __restrict__ spans,
	const ViewExt* __restrict__ views, uint32_t num_views, const uint32_t* __restrict__ band_lists,
	const uint32_t* __restrict__ band_counts, const typename CT::T* __restrict__ arena,
	const typename CT::T* __restrict__ skybox, typename CT::T* __restrict__ fb, uint32_t total_tiles,
	const uint32_t* __restrict__ counters)
{
	typedef typename CT::T C;
	__shared__ C tile[4][BAND_H][64];
	const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const uint32_t t = blockIdx.x * 4 + wave;
	if(t >= total_tiles || (counters[COUNTER_ERROR] & ERROR_SPAN_OVERFLOW))
		return;
	uint32_t lo = 0, hi = num_views;
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
	const int32_t y0 = int32_t(band * BAND_H);
	const int32_t W = int32_t(ve.v.width), H = int32_t(ve.v.height);
	const int32_t x = x0 + int32_t(lane);
	C* vfb = fb + ve.v.fb_offset;
	const C bg = C(ve.v.background_color);

#pragma unroll
	for(int r = 0; r < BAND_H; ++r)
	{
		const int32_t y = y0 + r;
		C c = bg;
		if(!ve.v.clear_background && x < W && y < H)
			c = vfb[uint64_t(y) * uint32_t(W) + uint32_t(x)];
		tile[wave][r][lane] = c;
	}

	const uint32_t band_global = ve.first_band + band;
	const uint32_t n = band_counts[band_global];
	const uint32_t* list = band_lists + ve.band_list_base + uint64_t(band) * ve.v.num_draws;
	const C* view_arena = arena + ve.v.surface_arena_offset;
	for(uint32_t e = 0; e < n; ++e)
	{
		const uint32_t di = list[e];
		const DrawRec dr = drec[di];
		if(dr.x_max <= x0 || dr.x_min >= x0 + 64)
			continue;
		const sw_draw& d = draws[di];
		const uint32_t blend = d.blend;
		const C* tex_c = d.texture_source == SW_TEXSRC_SKYBOX ? skybox + d.texture_offset : view_arena + d.texture_offset;
		const uint32_t* tex_lite = m.lite + d.texture_offset;
		for(uint32_t p = 0; p < dr.num_parts; ++p)
		{
			const PartRec pr = parts[di * MAX_PARTS_PER_DRAW + p];
			const int32_t ra = imax32(pr.row0, y0), rb = imin32(pr.row1, y0 + BAND_H);
			for(int32_t y = ra; y < rb; ++y)
			{
				const SpanRec s = spans[pr.span_base + uint32_t(y - pr.row0)];
				if(s.x1 <= x0 || s.x0 >= x0 + 64 || s.x0 >= s.x1)
					continue;
				if(x < s.x0 || x >= s.x1)
					continue;
				const int32_t j = x - s.x0;
				C& dst = tile[wave][y - y0][lane];
				if(pr.kind == 0)
				{
					const int32_t inv_z = wadd(s.v[0], wmul(j, s.v[1]));
					uint32_t div = uint32_t(inv_z) >> INV_Z_PRE_SHIFT;
					div = div ? div : 1u;
					const uint32_t z = (1u << Z_CALC_SHIFT) / div;
					const int32_t t0 = wadd(s.v[2], wmul(j, s.v[3]));
					const int32_t t1 = wadd(s.v[4], wmul(j, s.v[5]));
					const uint32_t u = uint32_t((uint64_t(z) * uint64_t(int64_t(t0))) >> TC_FINAL_SHIFT);
					const uint32_t v = uint32_t((uint64_t(z) * uint64_t(int64_t(t1))) >> TC_FINAL_SHIFT);
					bool ok;
					const C texel = fetch_checked<CT>(tex_c, u, v, d, ok);
					if(ok)
						dst = write_into_framebuffer<CT>(blend, dst, texel);
				}
				else if(pr.kind == 1 || pr.kind == 2)
				{
					const int32_t t0 = wadd(s.v[0], wmul(j, s.v[1]));
					const int32_t t1 = wadd(s.v[2], wmul(j, s.v[3]));
					const uint32_t u = uint32_t(pr.kind == 1 ? (t0 >> 16) : fixed16_floor_to_int(t0));
					const uint32_t v = uint32_t(pr.kind == 1 ? (t1 >> 16) : fixed16_floor_to_int(t1));
					bool ok;
					const C texel = fetch_checked<CT>(tex_c, u, v, d, ok);
					if(ok)
						dst = write_into_framebuffer<CT>(blend, dst, texel);
				}
				else
				{
					const int32_t t0 = wadd(s.v[0], wmul(j, pr.d[0]));
					const int32_t t1 = wadd(s.v[1], wmul(j, pr.d[1]));
					const int32_t u = fixed16_floor_to_int(t0), v = fixed16_floor_to_int(t1);
					if(uint32_t(u) < uint32_t(d.texture_size[0]) && uint32_t(v) < uint32_t(d.texture_size[1]))
					{
						const uint32_t texel = tex_lite[uint32_t(u) + uint32_t(v) * uint32_t(d.texture_size[0])];
						const int32_t lr = wadd(s.v[2], wmul(j, pr.d[2]));
						const int32_t lg = wadd(s.v[3], wmul(j, pr.d[3]));
						const int32_t lb = wadd(s.v[4], wmul(j, pr.d[4]));
						const CI tc = Color32T::to_ci(texel);
						const CI lit = CI{wmul(tc.b, lb) >> 16, wmul(tc.g, lg) >> 16, wmul(tc.r, lr) >> 16, 0};
						switch(blend)
						{
						case SW_BLEND_NONE: dst = CT::from_ci(lit); break;
						case SW_BLEND_AVERAGE: dst = CT::average(dst, CT::from_ci(lit)); break;
						case SW_BLEND_ADDITIVE: dst = CT::saturated_sum(dst, CT::from_ci(lit)); break;
						case SW_BLEND_ALPHA_TEST:
							if(Color32T::test_alpha(texel))
								dst = CT::from_ci(lit);
							break;
						case SW_BLEND_ALPHA_BLEND: dst = alpha_blend_ci<CT>(dst, lit, Color32T::get_alpha(texel)); break;
						default: break;
						}
					}
				}
			}
		}
	}

	if(x < W)
	{
#pragma unroll
		for(int r = 0; r < BAND_H; ++r)
		{
			const int32_t y = y0 + r;
			if(y < H)
				vfb[uint64_t(y) * uint32_t(W) + uint32_t(x)] = tile[wave][r][lane];
		}
	}
}

// =============================================================================================
// Tonemap (postprocessor.rs:237-253, :1019-1023)
// =============================================================================================
__global__ __launch_bounds__(256) void k_tonemap(const uint64_t* __restrict__ hdr, uint32_t width, uint32_t height,
	float inv_scale_r, float inv_scale_g, float inv_scale_b, uint32_t* __restrict__ dst, uint32_t dst_pitch,
	float* __restrict__ sums)
{
	__shared__ float red[3][256];
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	float sb = 0.0f, sg = 0.0f, sr = 0.0f;
	if(i < uint64_t(width) * height)
	{
		const uint32_t y = uint32_t(i / width), x = uint32_t(i - uint64_t(y) * width);
		const CV c = cv_from_color64(hdr[i]);
		sb = c.b;
		sg = c.g;
		sr = c.r;
		const float inv_255 = 1.0f / 255.0f;
		// div(c, mul_add(c, inv_255_vec, inv_scale_vec)); the alpha lane is c.a / fma(c.a, 0, 0)
		const CV mapped = CV{c.b / mul_add(c.b, inv_255, inv_scale_b), c.g / mul_add(c.g, inv_255, inv_scale_g),
			c.r / mul_add(c.r, inv_255, inv_scale_r), c.a / mul_add(c.a, 0.0f, 0.0f)};
		dst[uint64_t(y) * dst_pitch + x] = Color32T::from_cv(mapped);
	}
	red[0][threadIdx.x] = sr;
	red[1][threadIdx.x] = sg;
	red[2][threadIdx.x] = sb;
	__syncthreads();
	for(uint32_t s = 128; s > 0; s >>= 1)
	{
		if(threadIdx.x < s)
		{
			red[0][threadIdx.x] += red[0][threadIdx.x + s];
			red[1][threadIdx.x] += red[1][threadIdx.x + s];
			red[2][threadIdx.x] += red[2][threadIdx.x + s];
		}
		__syncthreads();
	}
	if(threadIdx.x == 0)
	{
		atomicAdd(&sums[0], red[0][0]);
		atomicAdd(&sums[1], red[1][0]);
		atomicAdd(&sums[2], red[2][0]);
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

void launch_prep_surfaces(hipStream_t st, const sw_surface_desc* s, uint32_t n, const ViewExt* views, uint32_t nv,
	uint64_t* surf_end, uint64_t arena_texels, const MapPtrs& m, uint32_t* counters)
{
	const uint32_t blocks = n ? (n + 255) / 256 : 1;
	hipLaunchKernelGGL(k_prep_surfaces, dim3(blocks), dim3(256), 0, st, s, n, views, nv, surf_end, arena_texels, m.num_levels,
		m.levels, m.num_lm_simple, m.num_lm_dir, counters);
}

void launch_build_surfaces(hipStream_t st, uint32_t color_bits, const MapPtrs& m, const sw_surface_desc* s, uint32_t n,
	const uint64_t* surf_end, const ViewExt* views, uint32_t nv, const sw_light* lights, uint32_t nl,
	const uint32_t* lidx, uint32_t nli, void* arena, uint64_t arena_texels, const uint32_t* counters)
{
	if(arena_texels == 0 || n == 0)
		return;
	const uint64_t blocks = (arena_texels + 255) / 256;
	if(color_bits == 64)
		hipLaunchKernelGGL(k_build_surfaces<Color64T>, dim3(uint32_t(blocks)), dim3(256), 0, st, m, s, n, surf_end, views, nv,
			lights, nl, lidx, nli, static_cast<uint64_t*>(arena), arena_texels, counters);
	else
		hipLaunchKernelGGL(k_build_surfaces<Color32T>, dim3(uint32_t(blocks)), dim3(256), 0, st, m, s, n, surf_end, views, nv,
			lights, nl, lidx, nli, static_cast<uint32_t*>(arena), arena_texels, counters);
}

void launch_setup_draws(hipStream_t st, const sw_draw* draws, uint32_t nd, const int32_t* vdata, uint64_t nvi,
	const ViewExt* views, uint32_t nv, PartRec* parts, SpanRec* spans, DrawRec* drec, uint32_t* counters, uint64_t span_capacity,
	uint64_t arena_texels, uint64_t skybox_texels, uint64_t lite_texels)
{
	if(nd == 0)
		return;
	hipLaunchKernelGGL(k_setup_draws, dim3((nd + 3) / 4), dim3(256), 0, st, draws, nd, vdata, nvi, views, nv, parts, spans, drec,
		counters, span_capacity, arena_texels, skybox_texels, lite_texels);
}

void launch_bin_bands(hipStream_t st, const DrawRec* drec, const ViewExt* views, uint32_t nv, uint32_t total_bands,
	uint32_t* band_lists, uint32_t* band_counts, const uint32_t* counters)
{
	if(total_bands == 0)
		return;
	hipLaunchKernelGGL(k_bin_bands, dim3(total_bands), dim3(256), 0, st, drec, views, nv, band_lists, band_counts, counters);
}

void launch_raster_tiles(hipStream_t st, uint32_t color_bits, const MapPtrs& m, const sw_draw* draws, const DrawRec* drec,
	const PartRec* parts, const SpanRec* spans, const ViewExt* views, uint32_t nv, const uint32_t* band_lists,
	const uint32_t* band_counts, const void* arena, const void* skybox, void* fb, uint32_t total_tiles,
	const uint32_t* counters)
{
	if(total_tiles == 0)
		return;
	const uint32_t blocks = (total_tiles + 3) / 4;
	if(color_bits == 64)
		hipLaunchKernelGGL(k_raster_tiles<Color64T>, dim3(blocks), dim3(256), 0, st, m, draws, drec, parts, spans, views, nv,
			band_lists, band_counts, static_cast<const uint64_t*>(arena), static_cast<const uint64_t*>(skybox),
			static_cast<uint64_t*>(fb), total_tiles, counters);
	else
		hipLaunchKernelGGL(k_raster_tiles<Color32T>, dim3(blocks), dim3(256), 0, st, m, draws, drec, parts, spans, views, nv,
			band_lists, band_counts, static_cast<const uint32_t*>(arena), static_cast<const uint32_t*>(skybox),
			static_cast<uint32_t*>(fb), total_tiles, counters);
}

void launch_tonemap(hipStream_t st, const uint64_t* hdr, uint32_t w, uint32_t h, float isr, float isg, float isb,
	uint32_t* dst, uint32_t dst_pitch, float* sums)
{
	hipLaunchKernelGGL(k_zero_floats, dim3(1), dim3(64), 0, st, sums, 3u);
	const uint64_t n = uint64_t(w) * h;
	if(n == 0)
		return;
	hipLaunchKernelGGL(k_tonemap, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, hdr, w, h, isr, isg, isb, dst, dst_pitch, sums);
}

} // namespace swk
