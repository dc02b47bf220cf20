// ORACLE -- CPU restatement of Square-Wheel's per-frame pixel path. TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
// only to check results against (or as the timed CPU baseline). The product path never calls it.
//
// It restates, loop for loop, the reference's Rust code:
//   - surface builder      square_wheel/surfaces.rs:433-852 (build_surface_impl_6_static_params),
//                          :854-971 (mix_surface_with_texture, also the portal textures' overlay of
//                          partial_renderer.rs:2205-2276), :973-1086 (lightmap element ops),
//                          :1088-1154 (specular / Fresnel helpers)
//   - span rasteriser      square_wheel/rasterizer.rs:137-271 (edge walk), :273-454 (perspective),
//                          :456-581 (line z corrected), :583-729 (affine), :765-1292 (triangles),
//                          :1511-1536 (write_into_framebuffer)
//   - colour model         square_wheel/fast_math.rs:6-368 (the SSE4.1 branch, with x86 intrinsics),
//                          square_wheel/abstract_color.rs:4-131, common/color.rs:2-168
//   - fixed point          common/fixed_math.rs:1-71
//   - tonemap              square_wheel/postprocessor.rs:226-256, :996-1024
//   - textures / mips      square_wheel/textures.rs:31-62, :93-161, :288-415
//   - screen split         square_wheel/rect_splitting.rs:23-78 (CPU baseline only)
//
// Two numeric modes: ORACLE_MODE_SSE uses rsqrtss / rcpss exactly as fast_math.rs:13-22 (host
// dependent, +-1 LSB vs. other hosts); ORACLE_MODE_EXACT uses 1/sqrt(x) and 1/x, the reference's
// own non-SSE fallback (fast_math.rs:376-384), and is host independent.
//
// Built with -msse4.1 -mfma -ffp-contract=off -fwrapv: FMA only where the reference calls
// f32_mul_add / _mm_fmadd_ps, every other float op unfused (Rust never contracts), integer
// overflow wraps as in a Rust release build.

#include "../include/sw_gpu.h"

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace
{

enum
{
	ORACLE_MODE_SSE = 0,
	ORACLE_MODE_EXACT = 1,
};

// ---------------------------------------------------------------------------------------------
// fast_math.rs
// ---------------------------------------------------------------------------------------------

struct Numerics
{
	int mode;

	// fast_math.rs:13-16 / :376-379
	float inv_sqrt_fast(float x) const
	{
		if(mode == ORACLE_MODE_SSE)
			return _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set1_ps(x)));
		return 1.0f / std::sqrt(x);
	}

	// fast_math.rs:19-22 / :381-384
	float inv_fast(float x) const
	{
		if(mode == ORACLE_MODE_SSE)
			return _mm_cvtss_f32(_mm_rcp_ss(_mm_set1_ps(x)));
		return 1.0f / x;
	}
};

// fast_math.rs:730-734 (release build has +fma)
inline float f32_mul_add(float x, float y, float z) { return std::fma(x, y, z); }

const float MIN_POSITIVE_VALUE = 1.0f / float(1 << 30); // fast_math.rs:744

// Rust f32::max / f32::min: IEEE maxNum / minNum (the non-NaN operand wins).
inline float fmax_r(float a, float b) { return std::fmax(a, b); }
inline float fmin_r(float a, float b) { return std::fmin(a, b); }

struct Vec3
{
	float x, y, z;
};

inline Vec3 v3(float x, float y, float z) { return Vec3{x, y, z}; }
inline Vec3 operator+(Vec3 a, Vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(Vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }

// fast_math.rs:780-783
inline float vec3_dot(Vec3 a, Vec3 b) { return f32_mul_add(a.x, b.x, f32_mul_add(a.y, b.y, a.z * b.z)); }
// fast_math.rs:786-793
inline Vec3 vec3_scalar_mul_add(Vec3 a, float s, Vec3 b)
{
	return v3(f32_mul_add(a.x, s, b.x), f32_mul_add(a.y, s, b.y), f32_mul_add(a.z, s, b.z));
}
inline float vec3_len2(Vec3 v) { return vec3_dot(v, v); }

// Shadow maps (light.rs:185-294), set by oracle_set_light_shadows for the following calls.
const sw_light_shadow* g_shadows = nullptr;
uint32_t g_num_shadows = 0;
const float* g_shadow_texels = nullptr;

// cube_shadow_map_side_fetch (light.rs:234-257)
float cube_shadow_map_side_fetch(const Numerics& nm, const sw_light_shadow& sm, Vec3 vec, uint32_t side)
{
	const float ONE_MINUS_EPS = 1.0f - 1.0f / 65536.0f;
	const float cubemap_size_f = float(sm.size);
	const float depth = nm.inv_fast(fmax_r(vec.z, MIN_POSITIVE_VALUE));
	const float half_depth = 0.5f * depth;
	const float u_f = fmin_r(fmax_r(f32_mul_add(vec.x, half_depth, 0.5f), 0.0f), ONE_MINUS_EPS) * cubemap_size_f;
	const float v_f = fmin_r(fmax_r(f32_mul_add(vec.y, half_depth, 0.5f), 0.0f), ONE_MINUS_EPS) * cubemap_size_f;
	const uint32_t u = uint32_t(u_f), v = uint32_t(v_f);
	const float value = g_shadow_texels[sm.offset + uint64_t(side) * sm.size * sm.size + u + uint64_t(v) * sm.size];
	return depth >= value ? 1.0f : 0.0f;
}

// get_light_shadow_factor (light.rs:185-193) with cube_shadow_map_fetch (:195-231) and
// projector_shadow_map_fetch (:259-294)
float get_light_shadow_factor(const Numerics& nm, uint32_t light_index, Vec3 vec)
{
	if(light_index >= g_num_shadows)
		return 1.0f;
	const sw_light_shadow& sm = g_shadows[light_index];
	if(sm.kind == SW_SHADOW_CUBE)
	{
		const Vec3 a = v3(std::fabs(vec.x), std::fabs(vec.y), std::fabs(vec.z));
		if(a.x >= a.y && a.x >= a.z)
			return vec.x >= 0.0f ? cube_shadow_map_side_fetch(nm, sm, v3(-vec.y, vec.z, a.x), 1)
								 : cube_shadow_map_side_fetch(nm, sm, v3(vec.y, vec.z, a.x), 0);
		if(a.y >= a.x && a.y >= a.z)
			return vec.y >= 0.0f ? cube_shadow_map_side_fetch(nm, sm, v3(vec.x, vec.z, a.y), 3)
								 : cube_shadow_map_side_fetch(nm, sm, v3(-vec.x, vec.z, a.y), 2);
		return vec.z >= 0.0f ? cube_shadow_map_side_fetch(nm, sm, v3(-vec.x, vec.y, a.z), 5)
							 : cube_shadow_map_side_fetch(nm, sm, v3(-vec.x, -vec.y, a.z), 4);
	}
	if(sm.kind == SW_SHADOW_PROJECTOR)
	{
		const Vec3 bx = v3(sm.basis_x[0], sm.basis_x[1], sm.basis_x[2]), by = v3(sm.basis_y[0], sm.basis_y[1], sm.basis_y[2]),
				   bz = v3(sm.basis_z[0], sm.basis_z[1], sm.basis_z[2]);
		const float z = vec3_dot(bz, vec);
		if(z <= 0.0f)
			return 0.0f;
		const float depth = nm.inv_fast(z);
		const float half_depth = 0.5f * depth;
		const float size_f = float(sm.size);
		const float u_f = f32_mul_add(vec3_dot(bx, vec), half_depth, 0.5f) * size_f;
		const float v_f = f32_mul_add(vec3_dot(by, vec), half_depth, 0.5f) * size_f;
		if(u_f < 0.0f || v_f < 0.0f || u_f >= size_f || v_f >= size_f)
			return 0.0f;
		const float value = g_shadow_texels[sm.offset + uint32_t(u_f) + uint64_t(uint32_t(v_f)) * sm.size];
		return depth >= value ? 1.0f : 0.0f;
	}
	return 1.0f;
}


// ColorVec (fast_math.rs:98-248): lanes [B, G, R, A] of an __m128.
inline __m128 cv_from_color32(uint32_t c) { return _mm_cvtepi32_ps(_mm_cvtepu8_epi32(_mm_cvtsi32_si128(int(c)))); }
inline __m128 cv_from_color64(uint64_t c)
{
	return _mm_cvtepi32_ps(_mm_cvtepu16_epi32(_mm_cvtsi64_si128((long long)c)));
}
// fast_math.rs:131-144: min(x, 255) -> cvtps2dq (round half even) -> low byte per lane.
inline uint32_t cv_into_color32(__m128 v)
{
	const __m128 clamped = _mm_min_ps(v, _mm_set1_ps(255.0f));
	const __m128i i = _mm_cvtps_epi32(clamped);
	const __m128i mask = _mm_set_epi8(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 12, 8, 4, 0);
	return uint32_t(_mm_cvtsi128_si32(_mm_shuffle_epi8(i, mask)));
}
// fast_math.rs:146-160
inline uint64_t cv_into_color64(__m128 v)
{
	const __m128 clamped = _mm_min_ps(v, _mm_set1_ps(65535.0f));
	const __m128i i = _mm_cvtps_epi32(clamped);
	const __m128i mask = _mm_set_epi8(0, 0, 0, 0, 0, 0, 0, 0, 13, 12, 9, 8, 5, 4, 1, 0);
	return uint64_t(_mm_cvtsi128_si64(_mm_shuffle_epi8(i, mask)));
}
inline __m128 cv_from_f32x3(const float c[3]) { return _mm_set_ps(0.0f, c[0], c[1], c[2]); }
inline __m128 cv_from_f32x3_with_one(const float c[3]) { return _mm_set_ps(1.0f, c[0], c[1], c[2]); }
inline __m128 cv_scalar_mul(__m128 a, float s) { return _mm_mul_ps(a, _mm_set1_ps(s)); }
inline __m128 cv_mul_scalar_add(__m128 a, float s, __m128 b) { return _mm_fmadd_ps(a, _mm_set1_ps(s), b); }
inline __m128 cv_insert3(__m128 a, float s) { return _mm_blend_ps(a, _mm_set1_ps(s), 1 << 3); }

// ColorVecI (fast_math.rs:258-367)
inline __m128i ci_from_color32(uint32_t c) { return _mm_cvtepu8_epi32(_mm_cvtsi32_si128(int(c))); }
inline __m128i ci_from_color64(uint64_t c) { return _mm_cvtepu16_epi32(_mm_cvtsi64_si128((long long)c)); }
inline __m128i ci_from_f32x3(const float c[3]) { return _mm_cvtps_epi32(_mm_set_ps(0.0f, c[0], c[1], c[2])); }
inline __m128i ci_from_i32x3(const int32_t c[3]) { return _mm_set_epi32(0, c[0], c[1], c[2]); }
inline uint32_t ci_into_color32(__m128i v)
{
	const __m128i zero = _mm_setzero_si128();
	return uint32_t(_mm_cvtsi128_si32(_mm_packus_epi16(_mm_packus_epi32(v, zero), zero)));
}
inline uint64_t ci_into_color64(__m128i v)
{
	return uint64_t(_mm_cvtsi128_si64(_mm_packus_epi32(v, _mm_setzero_si128())));
}

// AbstractColor (abstract_color.rs) for Color32 / Color64.
struct C32
{
	using T = uint32_t;
	static __m128 to_vec(T c) { return cv_from_color32(c); }
	static T from_vec(__m128 v) { return cv_into_color32(v); }
	static __m128i to_veci(T c) { return ci_from_color32(c); }
	static T from_veci(__m128i v) { return ci_into_color32(v); }
	// fast_math.rs:44-52 (pavgb)
	static T average(T a, T b)
	{
		return uint32_t(_mm_cvtsi128_si32(_mm_avg_epu8(_mm_cvtsi32_si128(int(a)), _mm_cvtsi32_si128(int(b)))));
	}
	// fast_math.rs:24-32 (paddusb)
	static T saturated_sum(T a, T b)
	{
		return uint32_t(_mm_cvtsi128_si32(_mm_adds_epu8(_mm_cvtsi32_si128(int(a)), _mm_cvtsi32_si128(int(b)))));
	}
	static int32_t get_alpha(T c) { return int32_t(c >> 24); }
	static bool test_alpha(T c) { return (c & 0xFF000000u) > 0x7F000000u; }
};

struct C64
{
	using T = uint64_t;
	static __m128 to_vec(T c) { return cv_from_color64(c); }
	static T from_vec(__m128 v) { return cv_into_color64(v); }
	static __m128i to_veci(T c) { return ci_from_color64(c); }
	static T from_veci(__m128i v) { return ci_into_color64(v); }
	static T average(T a, T b)
	{
		return uint64_t(
			_mm_cvtsi128_si64(_mm_avg_epu16(_mm_cvtsi64_si128((long long)a), _mm_cvtsi64_si128((long long)b))));
	}
	static T saturated_sum(T a, T b)
	{
		return uint64_t(
			_mm_cvtsi128_si64(_mm_adds_epu16(_mm_cvtsi64_si128((long long)a), _mm_cvtsi64_si128((long long)b))));
	}
	static int32_t get_alpha(T c) { return int32_t(c >> 48); }
	static bool test_alpha(T c) { return (c & 0xFFFF000000000000ull) > 0x007F000000000000ull; }
};

// abstract_color.rs:30-40 / :90-100
template <class CT>
typename CT::T alpha_blend(typename CT::T dst, typename CT::T src)
{
	const int32_t alpha = CT::get_alpha(src);
	const __m128i d = _mm_mullo_epi32(CT::to_veci(dst), _mm_set1_epi32(255 - alpha));
	const __m128i s = _mm_mullo_epi32(CT::to_veci(src), _mm_set1_epi32(alpha));
	return CT::from_veci(_mm_srai_epi32(_mm_add_epi32(d, s), 8));
}

// rasterizer.rs:1511-1536
template <class CT>
inline void write_into_framebuffer(uint32_t blend, typename CT::T& dst, typename CT::T texel)
{
	switch(blend)
	{
	case SW_BLEND_NONE: dst = texel; break;
	case SW_BLEND_AVERAGE: dst = CT::average(dst, texel); break;
	case SW_BLEND_ADDITIVE: dst = CT::saturated_sum(dst, texel); break;
	case SW_BLEND_ALPHA_TEST:
		if(CT::test_alpha(texel))
			dst = texel;
		break;
	case SW_BLEND_ALPHA_BLEND: dst = alpha_blend<CT>(dst, texel); break;
	default: break;
	}
}

// fast_math.rs:80-96 + textures.rs:41-61
inline void unpack_normal_roughness(int32_t packed, Vec3& n, float& r)
{
	const float s = 1.0f / 127.0f;
	const __m128 f = _mm_mul_ps(_mm_cvtepi32_ps(_mm_cvtepi8_epi32(_mm_cvtsi32_si128(packed))), _mm_set1_ps(s));
	alignas(16) float out[4];
	_mm_store_ps(out, f);
	n = v3(out[0], out[1], out[2]);
	r = out[3];
}
inline float unpack_roughness(int32_t packed) { return float(packed >> 24) * (1.0f / 127.0f); }

// fast_math.rs:64-78 + textures.rs:33-39
inline int32_t pack_normal_roughness(Vec3 n, float roughness)
{
	const __m128 v = _mm_set_ps(roughness, n.z, n.y, n.x);
	const __m128 scaled = _mm_mul_ps(v, _mm_set1_ps(127.0f));
	const __m128i i32v = _mm_cvtps_epi32(scaled);
	const __m128i zero = _mm_setzero_si128();
	return _mm_cvtsi128_si32(_mm_packs_epi16(_mm_packs_epi32(i32v, zero), zero));
}

// ---------------------------------------------------------------------------------------------
// fixed_math.rs
// ---------------------------------------------------------------------------------------------
typedef int32_t Fixed16;
const Fixed16 FIXED16_ONE = 1 << 16;
const Fixed16 FIXED16_HALF = FIXED16_ONE >> 1;

inline Fixed16 int_to_fixed16(int32_t x) { return int32_t(uint32_t(x) << 16); }
inline int32_t fixed16_round_to_int(Fixed16 x) { return (x + FIXED16_HALF) >> 16; }
inline int32_t fixed16_floor_to_int(Fixed16 x) { return x >> 16; }
inline float fixed16_to_f32(Fixed16 x) { return float(x) / float(FIXED16_ONE); }
inline Fixed16 f32_to_fixed16_unchecked(float x) { return _mm_cvttss_si32(_mm_set_ss(x * float(FIXED16_ONE))); }
inline Fixed16 fixed16_mul(Fixed16 x, Fixed16 y) { return int32_t((int64_t(x) * int64_t(y)) >> 16); }
inline Fixed16 fixed16_div(Fixed16 x, Fixed16 y) { return int32_t((int64_t(x) * 65536) / int64_t(y)); }
inline Fixed16 fixed16_mul_div(Fixed16 x, Fixed16 y, Fixed16 z) { return int32_t(int64_t(x) * int64_t(y) / int64_t(z)); }

// rasterizer.rs:1552-1556 (cvttss2si with 64-bit destination)
inline int64_t unchecked_to_int64(float x) { return _mm_cvttss_si64(_mm_set_ss(x)); }

const int TC_SHIFT = 24;
const int INV_Z_SHIFT = 29;
const int INV_Z_PRE_SHIFT = 8;
const int Z_CALC_SHIFT = 31;
const int TC_FINAL_SHIFT = TC_SHIFT + Z_CALC_SHIFT - INV_Z_SHIFT + INV_Z_PRE_SHIFT;
const float TC_SCALE = float(1 << TC_SHIFT);
const float INV_Z_SCALE = float(1 << INV_Z_SHIFT);

inline int32_t rem_euclid(int32_t a, int32_t b)
{
	const int32_t r = a % b;
	return r < 0 ? r + b : r;
}

// ---------------------------------------------------------------------------------------------
// Instrumentation (algorithmic work counts)
// ---------------------------------------------------------------------------------------------
// Work counters of one thread (the threaded path keeps one per thread and adds them after the join,
// so counting costs no shared-memory traffic in the timed CPU baseline).
struct Counters
{
	uint64_t surface_texels = 0, lightmap_elements = 0, emissive_texels = 0, polygon_pixels = 0, triangle_pixels = 0,
		blend_pixels = 0;
	// Texel fetches outside the texture. The reference debug_assert!s against them
	// (rasterizer.rs:436-437, :561-564, :709-712, :1231-1234); here they are skipped and counted.
	uint64_t out_of_bounds = 0;

	void add(const Counters& o)
	{
		surface_texels += o.surface_texels;
		lightmap_elements += o.lightmap_elements;
		emissive_texels += o.emissive_texels;
		polygon_pixels += o.polygon_pixels;
		triangle_pixels += o.triangle_pixels;
		blend_pixels += o.blend_pixels;
		out_of_bounds += o.out_of_bounds;
	}
};

inline bool blend_reads_dst(uint32_t blend)
{
	return blend == SW_BLEND_AVERAGE || blend == SW_BLEND_ADDITIVE || blend == SW_BLEND_ALPHA_BLEND ||
		blend == SW_BLEND_ALPHA_TEST;
}

// ---------------------------------------------------------------------------------------------
// surfaces.rs
// ---------------------------------------------------------------------------------------------
struct MapData
{
	const sw_texel* texels;
	const uint32_t* lite_texels;
	const float* lightmaps_simple; // [f32; 3] per element
	const sw_directional_lightmap_element* lightmaps_directional;
	const sw_texture_level* levels;
	const void* skybox;
};

struct LmElem
{
	float ambient[3];
	Vec3 dir;
	float deviation;
	float color[3];
};

// surfaces.rs:996-1004 / :1022-1070: mix(a, b, ratio) = fma(a, ratio, b * (1 - ratio))
inline LmElem lm_mix(const LmElem& a, const LmElem& b, float ratio, bool directional)
{
	const float omr = 1.0f - ratio;
	LmElem r;
	for(int i = 0; i < 3; ++i)
		r.ambient[i] = f32_mul_add(a.ambient[i], ratio, b.ambient[i] * omr);
	if(directional)
	{
		r.dir = v3(
			f32_mul_add(a.dir.x, ratio, b.dir.x * omr), f32_mul_add(a.dir.y, ratio, b.dir.y * omr),
			f32_mul_add(a.dir.z, ratio, b.dir.z * omr));
		r.deviation = f32_mul_add(a.deviation, ratio, b.deviation * omr);
		for(int i = 0; i < 3; ++i)
			r.color[i] = f32_mul_add(a.color[i], ratio, b.color[i] * omr);
	}
	else
	{
		r.dir = v3(0, 0, 0);
		r.deviation = 0.0f;
		r.color[0] = r.color[1] = r.color[2] = 0.0f;
	}
	return r;
}

inline LmElem lm_load(const MapData& m, bool directional, uint64_t index)
{
	LmElem e;
	if(directional)
	{
		const sw_directional_lightmap_element& d = m.lightmaps_directional[index];
		for(int i = 0; i < 3; ++i)
		{
			e.ambient[i] = d.ambient_light[i];
			e.color[i] = d.directional_light_color[i];
		}
		e.dir = v3(d.light_direction_vector_scaled[0], d.light_direction_vector_scaled[1], d.light_direction_vector_scaled[2]);
		e.deviation = d.directional_light_deviation;
	}
	else
	{
		for(int i = 0; i < 3; ++i)
			e.ambient[i] = m.lightmaps_simple[index * 3 + i];
		e.dir = v3(0, 0, 0);
		e.deviation = 0.0f;
		e.color[0] = e.color[1] = e.color[2] = 0.0f;
	}
	return e;
}

const uint32_t SPECULAR_TYPE_NONE = 0, SPECULAR_TYPE_DIELECTRIC = 1, SPECULAR_TYPE_METAL = 2;
const float DIELECTRIC_ZERO_REFLECTIVITY = 0.04f;
const float DIELECTRIC_AVERAGE_REFLECTIVITY = DIELECTRIC_ZERO_REFLECTIVITY * 3.0f;
const float METAL_AVERAGE_SCHLICK_FACTOR = 0.5f;

// surfaces.rs:1088-1114 (the "new formula" branch)
inline float get_specular_intensity(const Numerics& nm, float angle_cos, float inv_roughness)
{
	const float x = (angle_cos - 1.0f) * inv_roughness;
	const float SQRT_7 = 2.645751311f;
	const float A = 2.0f * float(M_PI) / SQRT_7;
	const float B = float(M_PI) * SQRT_7 / 2.0f;
	return nm.inv_fast(f32_mul_add(x * x, B, A)) * inv_roughness;
}
// surfaces.rs:1116-1123
inline float get_fresnel_factor_base(float angle_cos)
{
	const float a = fmax_r(1.0f - angle_cos, 0.0f);
	const float a2 = a * a;
	return a2 * a2 * a;
}
// surfaces.rs:1125-1141
inline float get_specular_k_dielectric(float f_base, float roughness)
{
	const float f = f32_mul_add(f_base, 1.0f - DIELECTRIC_ZERO_REFLECTIVITY, DIELECTRIC_ZERO_REFLECTIVITY);
	return f32_mul_add(f, 1.0f - roughness, DIELECTRIC_AVERAGE_REFLECTIVITY * roughness);
}
// surfaces.rs:1143-1150
inline float get_specular_k_metal(float f_base, float roughness)
{
	return f32_mul_add(f_base, 1.0f - roughness, METAL_AVERAGE_SCHLICK_FACTOR * roughness);
}

// surfaces.rs:433-852 (all static specialisations folded into runtime flags).
template <class CT>
void build_surface(
	const Numerics& nm, const MapData& m, const sw_surface_desc& s, const sw_light* lights,
	const uint32_t* light_indices, Vec3 cam_pos, typename CT::T* out, Counters* counters)
{
	const sw_texture_level& tex = m.levels[s.texture_level];
	const sw_texel* tex_pixels = m.texels + tex.offset;
	const bool use_normal_map = (tex.flags & SW_TEX_HAS_NORMAL_MAP) != 0;
	const uint32_t specular_type = (tex.flags & SW_TEX_HAS_NON_ONE_ROUGHNESS) ?
		((tex.flags & SW_TEX_IS_METAL) ? SPECULAR_TYPE_METAL : SPECULAR_TYPE_DIELECTRIC) :
		SPECULAR_TYPE_NONE;
	const bool has_lightmap = s.lightmap_offset != 0;
	const bool directional = (s.flags & SW_SURF_DIRECTIONAL) != 0;
	const uint32_t num_lights = std::min<uint32_t>(s.num_lights, SW_MAX_POLYGON_LIGHTS);
	const bool use_dynamic_lights = num_lights > 0;
	const uint32_t scale_log2 = s.lightmap_scale_log2;

	const Vec3 u_vec = v3(s.basis_u[0], s.basis_u[1], s.basis_u[2]);
	const Vec3 v_vec = v3(s.basis_v[0], s.basis_v[1], s.basis_v[2]);
	const Vec3 basis_start = v3(s.basis_start[0], s.basis_start[1], s.basis_start[2]);
	const Vec3 basis_normal = v3(s.basis_normal[0], s.basis_normal[1], s.basis_normal[2]);

	// :456-457
	const Vec3 start_pos =
		basis_start + u_vec * (float(s.tc_start[0]) + 0.5f) + v_vec * (float(s.tc_start[1]) + 0.5f);
	// :461-462
	const Vec3 u_vec_normalized = u_vec * nm.inv_sqrt_fast(fmax_r(vec3_len2(u_vec), MIN_POSITIVE_VALUE));
	const Vec3 v_vec_normalized = v_vec * nm.inv_sqrt_fast(fmax_r(vec3_len2(v_vec), MIN_POSITIVE_VALUE));

	LmElem line_lightmap[SW_MAX_LIGHTMAP_SIZE + 2];
	std::memset(line_lightmap, 0, sizeof(line_lightmap));
	const float lightmap_scale_f = float(1u << (has_lightmap ? scale_log2 : 0));
	const float inv_lightmap_scale_f = 1.0f / lightmap_scale_f;
	const float k_shift = 0.5f * inv_lightmap_scale_f;
	const uint32_t lightmap_fetch_mask = (1u << (has_lightmap ? scale_log2 : 0)) - 1u;

	const uint32_t w = s.size[0], h = s.size[1];
	for(uint32_t dst_v = 0; dst_v < h; ++dst_v)
	{
		if(has_lightmap)
		{
			// :477-502
			const uint32_t lightmap_base_v = dst_v + s.lightmap_tc_shift[1];
			const uint32_t lightmap_v = lightmap_base_v >> scale_log2;
			const float k = f32_mul_add(float(lightmap_base_v & lightmap_fetch_mask), inv_lightmap_scale_f, k_shift);
			const uint64_t base = uint64_t(s.lightmap_offset) + uint64_t(lightmap_v) * s.lightmap_size[0];
			const uint32_t n = std::min<uint32_t>(s.lightmap_size[0], SW_MAX_LIGHTMAP_SIZE + 2);
			for(uint32_t i = 0; i < n; ++i)
			{
				const LmElem l0 = lm_load(m, directional, base + i);
				const LmElem l1 = lm_load(m, directional, base + s.lightmap_size[0] + i);
				line_lightmap[i] = lm_mix(l1, l0, k, directional);
			}
			if(counters)
				counters->lightmap_elements += 2 * uint64_t(s.lightmap_size[0]);
		}

		typename CT::T* dst_line = out + uint64_t(dst_v) * w;
		const int32_t src_v = rem_euclid(s.tc_start[1] + int32_t(dst_v), int32_t(tex.height));
		const sw_texel* src_line = tex_pixels + uint64_t(uint32_t(src_v)) * tex.width;
		int32_t src_u = rem_euclid(s.tc_start[0], int32_t(tex.width));
		const Vec3 start_pos_v = vec3_scalar_mul_add(v_vec, float(dst_v), start_pos);

		for(uint32_t dst_u = 0; dst_u < w; ++dst_u)
		{
			const Vec3 pos = vec3_scalar_mul_add(u_vec, float(dst_u), start_pos_v);
			const sw_texel texel_value = src_line[src_u];

			Vec3 texel_normal;
			float texel_roughness;
			if(use_normal_map)
				unpack_normal_roughness(texel_value.packed_normal_roughness, texel_normal, texel_roughness);
			else if(specular_type == SPECULAR_TYPE_NONE)
			{
				texel_normal = v3(0, 0, 1);
				texel_roughness = 0.0f;
			}
			else
			{
				texel_normal = v3(0, 0, 1);
				texel_roughness = unpack_roughness(texel_value.packed_normal_roughness);
			}

			__m128 total_light_albedo_modulated = _mm_setzero_ps();
			__m128 total_light_direct = _mm_setzero_ps();
			if(has_lightmap)
			{
				// :538-549
				const uint32_t lightmap_base_u = dst_u + s.lightmap_tc_shift[0];
				const uint32_t lightmap_u = lightmap_base_u >> scale_log2;
				const LmElem& l0 = line_lightmap[lightmap_u];
				const LmElem& l1 = line_lightmap[lightmap_u + 1];
				const float k = f32_mul_add(float(lightmap_base_u & lightmap_fetch_mask), inv_lightmap_scale_f, k_shift);
				const LmElem l_mixed = lm_mix(l1, l0, k, directional);

				if(specular_type == SPECULAR_TYPE_NONE)
				{
					// :551-572
					total_light_albedo_modulated = cv_from_f32x3_with_one(l_mixed.ambient);
					if(directional)
					{
						const float dot =
							use_normal_map ? fmax_r(vec3_dot(l_mixed.dir, texel_normal), 0.0f) : l_mixed.dir.z;
						total_light_albedo_modulated =
							cv_mul_scalar_add(cv_from_f32x3(l_mixed.color), dot, total_light_albedo_modulated);
					}
				}
				else
				{
					// :575-677
					const Vec3 normal = use_normal_map ? texel_normal : v3(0, 0, 1);
					const Vec3 vec_to_camera = cam_pos - pos;
					const Vec3 vec_to_camera_ts = v3(
						vec3_dot(vec_to_camera, u_vec_normalized), vec3_dot(vec_to_camera, v_vec_normalized),
						vec3_dot(vec_to_camera, basis_normal));
					const float vec_to_camera_normal_dot = vec3_dot(vec_to_camera_ts, normal);
					const Vec3 vec_to_camera_reflected = normal * (2.0f * vec_to_camera_normal_dot) - vec_to_camera_ts;
					const float vec_to_camera_len2 = fmax_r(vec3_len2(vec_to_camera_reflected), MIN_POSITIVE_VALUE);
					const float angle_cos =
						fmax_r(vec_to_camera_normal_dot * nm.inv_sqrt_fast(vec_to_camera_len2), 0.0f);
					const float fresnel_factor_base = get_fresnel_factor_base(angle_cos);
					const float specular_k = specular_type == SPECULAR_TYPE_DIELECTRIC ?
						get_specular_k_dielectric(fresnel_factor_base, texel_roughness) :
						get_specular_k_metal(fresnel_factor_base, texel_roughness);
					const float one_minus_specular_k = 1.0f - specular_k;
					const __m128 constant_component = cv_from_f32x3(l_mixed.ambient);
					total_light_albedo_modulated = cv_scalar_mul(constant_component, one_minus_specular_k);
					total_light_direct = cv_scalar_mul(constant_component, specular_k);
					total_light_albedo_modulated = cv_insert3(total_light_albedo_modulated, 1.0f);

					if(directional)
					{
						const float direction_vec_len2 = vec3_len2(l_mixed.dir);
						const float direction_vec_len = direction_vec_len2 * nm.inv_sqrt_fast(direction_vec_len2);
						const float reflected_light_angle_cos = vec3_dot(vec_to_camera_reflected, l_mixed.dir) *
							nm.inv_sqrt_fast(vec_to_camera_len2 * direction_vec_len2);
						const float inv_roughness_corrected =
							fmax_r(nm.inv_fast(texel_roughness + l_mixed.deviation), 0.75f);
						const float specular_intensity =
							get_specular_intensity(nm, reflected_light_angle_cos, inv_roughness_corrected);
						const __m128 dir_color = cv_from_f32x3(l_mixed.color);
						if(specular_type == SPECULAR_TYPE_DIELECTRIC)
						{
							const float diffuse_intensity = fmax_r(vec3_dot(l_mixed.dir, texel_normal), 0.0f);
							const float light_intensity_diffuse = diffuse_intensity * one_minus_specular_k;
							total_light_albedo_modulated =
								cv_mul_scalar_add(dir_color, light_intensity_diffuse, total_light_albedo_modulated);
							const float light_intensity_specular = specular_intensity * specular_k * direction_vec_len;
							total_light_direct = cv_mul_scalar_add(dir_color, light_intensity_specular, total_light_direct);
						}
						else
						{
							const float scale_factor = specular_intensity * direction_vec_len;
							const float light_intensity_modulated = one_minus_specular_k * scale_factor;
							total_light_albedo_modulated =
								cv_mul_scalar_add(dir_color, light_intensity_modulated, total_light_albedo_modulated);
							const float light_intensity_direct = specular_k * scale_factor;
							total_light_direct = cv_mul_scalar_add(dir_color, light_intensity_direct, total_light_direct);
						}
					}
				}
			}

			if(use_dynamic_lights)
			{
				// :683-707
				Vec3 normal;
				if(use_normal_map)
				{
					normal = v3(
						f32_mul_add(
							texel_normal.x, u_vec_normalized.x,
							f32_mul_add(texel_normal.y, v_vec_normalized.x, texel_normal.z * basis_normal.x)),
						f32_mul_add(
							texel_normal.x, u_vec_normalized.y,
							f32_mul_add(texel_normal.y, v_vec_normalized.y, texel_normal.z * basis_normal.y)),
						f32_mul_add(
							texel_normal.x, u_vec_normalized.z,
							f32_mul_add(texel_normal.y, v_vec_normalized.z, texel_normal.z * basis_normal.z)));
				}
				else
					normal = basis_normal;

				Vec3 vec_to_camera_reflected;
				float vec_to_camera_len2, specular_k, inv_roughness;
				if(specular_type != SPECULAR_TYPE_NONE)
				{
					// :714-741
					const Vec3 vec_to_camera = cam_pos - pos;
					const float vec_to_camera_normal_dot = vec3_dot(vec_to_camera, normal);
					vec_to_camera_reflected = normal * (2.0f * vec_to_camera_normal_dot) - vec_to_camera;
					vec_to_camera_len2 = fmax_r(vec3_len2(vec_to_camera_reflected), MIN_POSITIVE_VALUE);
					const float angle_cos =
						fmax_r(vec_to_camera_normal_dot * nm.inv_sqrt_fast(vec_to_camera_len2), 0.0f);
					const float fresnel_factor_base = get_fresnel_factor_base(angle_cos);
					specular_k = specular_type == SPECULAR_TYPE_DIELECTRIC ?
						get_specular_k_dielectric(fresnel_factor_base, texel_roughness) :
						get_specular_k_metal(fresnel_factor_base, texel_roughness);
					inv_roughness = nm.inv_fast(texel_roughness);
				}
				else
				{
					vec_to_camera_reflected = v3(0, 0, 0);
					vec_to_camera_len2 = MIN_POSITIVE_VALUE;
					specular_k = 0.0f;
					inv_roughness = 0.0f;
				}

				for(uint32_t li = 0; li < num_lights; ++li)
				{
					// :750-830
					const uint32_t light_index = light_indices[s.first_light + li];
					const sw_light& light = lights[light_index];
					const Vec3 light_pos = v3(light.position[0], light.position[1], light.position[2]);
					const Vec3 vec_to_light = light_pos - pos;
					const float shadow_factor = get_light_shadow_factor(nm, light_index, vec_to_light);
					const float vec_to_light_len2 = fmax_r(vec3_len2(vec_to_light), MIN_POSITIVE_VALUE);
					const float distance_factor = fmax_r(nm.inv_fast(vec_to_light_len2) - light.inv_square_radius, 0.0f);
					const float shadow_distance_factor = shadow_factor * distance_factor;
					const float diffuse_intensity = specular_type == SPECULAR_TYPE_METAL ?
						0.0f :
						fmax_r(vec3_dot(normal, vec_to_light) * nm.inv_sqrt_fast(vec_to_light_len2), 0.0f);
					float specular_intensity = 0.0f;
					if(specular_type != SPECULAR_TYPE_NONE)
					{
						const float c = vec3_dot(vec_to_camera_reflected, vec_to_light) *
							nm.inv_sqrt_fast(vec_to_camera_len2 * vec_to_light_len2);
						specular_intensity = get_specular_intensity(nm, c, inv_roughness);
					}
					const __m128 light_color = cv_from_f32x3(light.color);
					if(specular_type == SPECULAR_TYPE_NONE)
					{
						total_light_albedo_modulated = cv_mul_scalar_add(
							light_color, diffuse_intensity * shadow_distance_factor, total_light_albedo_modulated);
					}
					else if(specular_type == SPECULAR_TYPE_DIELECTRIC)
					{
						const float light_intensity_diffuse =
							diffuse_intensity * (1.0f - specular_k) * shadow_distance_factor;
						total_light_albedo_modulated =
							cv_mul_scalar_add(light_color, light_intensity_diffuse, total_light_albedo_modulated);
						const float light_intensity_specular = specular_intensity * specular_k * shadow_distance_factor;
						total_light_direct = cv_mul_scalar_add(light_color, light_intensity_specular, total_light_direct);
					}
					else
					{
						const float sisdf = specular_intensity * shadow_distance_factor;
						const float light_intensity_modulated = (1.0f - specular_k) * sisdf;
						total_light_albedo_modulated =
							cv_mul_scalar_add(light_color, light_intensity_modulated, total_light_albedo_modulated);
						const float light_intensity_direct = specular_k * sisdf;
						total_light_direct = cv_mul_scalar_add(light_color, light_intensity_direct, total_light_direct);
					}
				}
			}

			// :833-842
			__m128 result_color = _mm_mul_ps(cv_from_color32(texel_value.diffuse), total_light_albedo_modulated);
			if(specular_type != SPECULAR_TYPE_NONE)
				result_color = cv_mul_scalar_add(total_light_direct, 255.0f, result_color);
			dst_line[dst_u] = CT::from_vec(result_color);

			++src_u;
			if(src_u == int32_t(tex.width))
				src_u = 0;
		}
	}
	if(counters)
		counters->surface_texels += uint64_t(w) * h;
}

// surfaces.rs:903-971
template <class CT>
void mix_surface_with_texture(
	const MapData& m, const sw_surface_desc& s, uint32_t blend, typename CT::T* data, Counters* counters)
{
	const sw_texture_level& tex = m.levels[s.emissive_level];
	const uint32_t* tex_pixels = m.lite_texels + tex.offset;
	const float light_scale = float(1 << 8);
	const float lv[3] = {s.emissive_light[0] * light_scale, s.emissive_light[1] * light_scale, s.emissive_light[2] * light_scale};
	const __m128i light_vec = ci_from_f32x3(lv);
	for(uint32_t dst_v = 0; dst_v < s.size[1]; ++dst_v)
	{
		typename CT::T* dst_line = data + uint64_t(dst_v) * s.size[0];
		const int32_t src_v = rem_euclid(s.tc_start[1] + int32_t(dst_v), int32_t(tex.height));
		const uint32_t* src_line = tex_pixels + uint64_t(uint32_t(src_v)) * tex.width;
		int32_t src_u = rem_euclid(s.tc_start[0], int32_t(tex.width));
		for(uint32_t x = 0; x < s.size[0]; ++x)
		{
			const uint32_t texel_value = src_line[src_u];
			const __m128i modulated = _mm_srai_epi32(_mm_mullo_epi32(ci_from_color32(texel_value), light_vec), 8);
			typename CT::T& dst = dst_line[x];
			switch(blend)
			{
			case SW_BLEND_NONE: dst = CT::from_veci(modulated); break;
			case SW_BLEND_AVERAGE:
				dst = CT::from_veci(_mm_srai_epi32(_mm_add_epi32(modulated, CT::to_veci(dst)), 1));
				break;
			case SW_BLEND_ADDITIVE: dst = CT::from_veci(_mm_add_epi32(modulated, CT::to_veci(dst))); break;
			case SW_BLEND_ALPHA_TEST:
				if(C32::test_alpha(texel_value))
					dst = CT::from_veci(modulated);
				break;
			case SW_BLEND_ALPHA_BLEND:
			{
				const int32_t alpha = C32::get_alpha(texel_value);
				dst = CT::from_veci(_mm_srai_epi32(
					_mm_add_epi32(
						_mm_mullo_epi32(modulated, _mm_set1_epi32(alpha)),
						_mm_mullo_epi32(CT::to_veci(dst), _mm_set1_epi32(255 - alpha))),
					8));
				break;
			}
			default: break;
			}
			++src_u;
			if(src_u == int32_t(tex.width))
				src_u = 0;
		}
	}
	if(counters)
		counters->emissive_texels += uint64_t(s.size[0]) * s.size[1];
}

// ---------------------------------------------------------------------------------------------
// rasterizer.rs
// ---------------------------------------------------------------------------------------------
struct ClipRect
{
	int32_t min_x, min_y, max_x, max_y;
};

struct PolygonSide
{
	Fixed16 x_start, dx_dy;
};

template <class CT>
struct Rasterizer
{
	typename CT::T* color_buffer;
	int32_t row_size;
	ClipRect clip_rect;
	Counters* counters;

	bool in_texture(uint32_t u, uint32_t v, const int32_t* tsize)
	{
		if(u < uint32_t(tsize[0]) && v < uint32_t(tsize[1]))
			return true;
		if(counters)
			counters->out_of_bounds++;
		return false;
	}

	void count(uint32_t blend, int64_t n, bool triangle)
	{
		if(!counters || n <= 0)
			return;
		if(triangle)
			counters->triangle_pixels += uint64_t(n);
		else
			counters->polygon_pixels += uint64_t(n);
		if(blend_reads_dst(blend))
			counters->blend_pixels += uint64_t(n);
	}

	// rasterizer.rs:273-454
	void fill_polygon_part(
		uint32_t blend, Fixed16 y_start, Fixed16 y_end, PolygonSide left_side, PolygonSide right_side, const float* de,
		const float* te, const int32_t* tsize, const typename CT::T* texture_data)
	{
		const int64_t d_inv_z_dx = unchecked_to_int64(INV_Z_SCALE * de[0]);
		const int64_t d_inv_z_dy = unchecked_to_int64(INV_Z_SCALE * de[1]);
		const int64_t inv_z_k = unchecked_to_int64(INV_Z_SCALE * (de[2] + (de[0] + de[1]) * 0.5f));
		const int64_t d_tc_dx[2] = {unchecked_to_int64(TC_SCALE * te[0]), unchecked_to_int64(TC_SCALE * te[1])};
		const int64_t d_tc_dy[2] = {unchecked_to_int64(TC_SCALE * te[2]), unchecked_to_int64(TC_SCALE * te[3])};
		const int64_t tc_k[2] = {
			unchecked_to_int64(TC_SCALE * (te[4] + (te[0] + te[2]) * 0.5f)),
			unchecked_to_int64(TC_SCALE * (te[5] + (te[1] + te[3]) * 0.5f))};
		const uint32_t texture_width = uint32_t(tsize[0]);

		const int32_t y_start_int = std::max(fixed16_round_to_int(y_start), clip_rect.min_y);
		const int32_t y_end_int = std::min(fixed16_round_to_int(y_end), clip_rect.max_y);
		const Fixed16 y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y_start;
		Fixed16 x_left = left_side.x_start + fixed16_mul(y_start_delta, left_side.dx_dy) + FIXED16_HALF;
		Fixed16 x_right = right_side.x_start + fixed16_mul(y_start_delta, right_side.dx_dy) + FIXED16_HALF;
		int64_t line_inv_z = int64_t(y_start_int) * d_inv_z_dy + inv_z_k;
		int64_t line_tc[2] = {int64_t(y_start_int) * d_tc_dy[0] + tc_k[0], int64_t(y_start_int) * d_tc_dy[1] + tc_k[1]};

		for(int32_t y_int = y_start_int; y_int < y_end_int; ++y_int)
		{
			const int32_t x_start_int = std::max(fixed16_floor_to_int(x_left), clip_rect.min_x);
			const int32_t x_end_int = std::min(fixed16_floor_to_int(x_right), clip_rect.max_x);
			if(x_start_int < x_end_int)
			{
				typename CT::T* line_dst = color_buffer + int64_t(y_int) * row_size;
				const int64_t span_start_x = x_start_int;
				const int64_t span_end_x = x_end_int - 1;
				const int32_t span_length_minus_one = (x_end_int - 1) - x_start_int;

				const int64_t inv_z_min = 1 << INV_Z_PRE_SHIFT;
				const int64_t inv_z_max = 1 << 29;
				const int32_t span_inv_z_start =
					int32_t(std::max(inv_z_min, std::min(span_start_x * d_inv_z_dx + line_inv_z, inv_z_max)));
				const int32_t span_inv_z_end =
					int32_t(std::max(inv_z_min, std::min(span_end_x * d_inv_z_dx + line_inv_z, inv_z_max)));
				const int32_t span_d_inv_z =
					span_length_minus_one > 0 ? (span_inv_z_end - span_inv_z_start) / span_length_minus_one : 0;
				int32_t span_inv_z = span_inv_z_start;

				int32_t span_tc[2] = {0, 0}, span_d_tc[2] = {0, 0};
				for(int i = 0; i < 2; ++i)
				{
					const int tc_max_shift = INV_Z_SHIFT - TC_SHIFT;
					const int64_t tc_max_max = 1 << 29;
					const int64_t span_inv_z_corrected = int64_t(span_inv_z - (1 << INV_Z_PRE_SHIFT));
					const int64_t tc_max_start = std::max<int64_t>(
						0, std::min(((span_inv_z_corrected * int64_t(tsize[i])) >> tc_max_shift) - 1, tc_max_max));
					const int32_t span_tc_start =
						int32_t(std::max<int64_t>(0, std::min(span_start_x * d_tc_dx[i] + line_tc[i], tc_max_start)));
					if(span_length_minus_one > 0)
					{
						const int64_t tc_max_end = std::max<int64_t>(
							0,
							std::min(
								(((span_inv_z_corrected + int64_t(int32_t(span_d_inv_z * span_length_minus_one))) *
								  int64_t(tsize[i])) >>
								 tc_max_shift) -
									1,
								tc_max_max));
						const int32_t span_tc_end =
							int32_t(std::max<int64_t>(0, std::min(span_end_x * d_tc_dx[i] + line_tc[i], tc_max_end)));
						span_d_tc[i] = (span_tc_end - span_tc_start) / span_length_minus_one;
						if(span_tc_end > span_tc_start)
							span_tc[i] = span_tc_start;
						else
							span_tc[i] = span_tc_end - span_d_tc[i] * span_length_minus_one;
					}
					else
						span_tc[i] = span_tc_start;
				}

				for(int32_t x = x_start_int; x < x_end_int; ++x)
				{
					const uint32_t z = (1u << Z_CALC_SHIFT) / (uint32_t(span_inv_z) >> INV_Z_PRE_SHIFT);
					const uint32_t pix_tc0 = uint32_t((uint64_t(z) * uint64_t(int64_t(span_tc[0]))) >> TC_FINAL_SHIFT);
					const uint32_t pix_tc1 = uint32_t((uint64_t(z) * uint64_t(int64_t(span_tc[1]))) >> TC_FINAL_SHIFT);
					if(in_texture(pix_tc0, pix_tc1, tsize))
						write_into_framebuffer<CT>(blend, line_dst[x], texture_data[pix_tc0 + pix_tc1 * texture_width]);
					span_inv_z += span_d_inv_z;
					span_tc[0] += span_d_tc[0];
					span_tc[1] += span_d_tc[1];
				}
				count(blend, x_end_int - x_start_int, false);
			}
			x_left += left_side.dx_dy;
			x_right += right_side.dx_dy;
			line_inv_z += d_inv_z_dy;
			line_tc[0] += d_tc_dy[0];
			line_tc[1] += d_tc_dy[1];
		}
	}

	// rasterizer.rs:456-581
	void fill_polygon_part_line_z_corrected(
		uint32_t blend, Fixed16 y_start, Fixed16 y_end, PolygonSide left_side, PolygonSide right_side, const float* de,
		const float* te, const int32_t* tsize, const typename CT::T* texture_data)
	{
		const int LINE_TC_SHIFT = 16;
		const int64_t d_inv_z_dx = unchecked_to_int64(INV_Z_SCALE * de[0]);
		const int64_t d_inv_z_dy = unchecked_to_int64(INV_Z_SCALE * de[1]);
		const int64_t inv_z_k = unchecked_to_int64(INV_Z_SCALE * (de[2] + (de[0] + de[1]) * 0.5f));
		const int64_t d_tc_dx[2] = {unchecked_to_int64(TC_SCALE * te[0]), unchecked_to_int64(TC_SCALE * te[1])};
		const int64_t d_tc_dy[2] = {unchecked_to_int64(TC_SCALE * te[2]), unchecked_to_int64(TC_SCALE * te[3])};
		const int64_t tc_k[2] = {
			unchecked_to_int64(TC_SCALE * (te[4] + (te[0] + te[2]) * 0.5f)),
			unchecked_to_int64(TC_SCALE * (te[5] + (te[1] + te[3]) * 0.5f))};
		const uint32_t texture_width = uint32_t(tsize[0]);

		const int32_t y_start_int = std::max(fixed16_round_to_int(y_start), clip_rect.min_y);
		const int32_t y_end_int = std::min(fixed16_round_to_int(y_end), clip_rect.max_y);
		const Fixed16 y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y_start;
		Fixed16 x_left = left_side.x_start + fixed16_mul(y_start_delta, left_side.dx_dy) + FIXED16_HALF;
		Fixed16 x_right = right_side.x_start + fixed16_mul(y_start_delta, right_side.dx_dy) + FIXED16_HALF;
		int64_t line_inv_z = int64_t(y_start_int) * d_inv_z_dy + inv_z_k;
		int64_t line_tc[2] = {int64_t(y_start_int) * d_tc_dy[0] + tc_k[0], int64_t(y_start_int) * d_tc_dy[1] + tc_k[1]};

		for(int32_t y_int = y_start_int; y_int < y_end_int; ++y_int)
		{
			const int32_t x_start_int = std::max(fixed16_floor_to_int(x_left), clip_rect.min_x);
			const int32_t x_end_int = std::min(fixed16_floor_to_int(x_right), clip_rect.max_x);
			if(x_start_int < x_end_int)
			{
				typename CT::T* line_dst = color_buffer + int64_t(y_int) * row_size;
				const int64_t span_start_x = x_start_int;
				const int64_t span_end_x = x_end_int - 1;
				const int32_t span_length_minus_one = (x_end_int - 1) - x_start_int;
				const int64_t inv_z_min = 1 << INV_Z_PRE_SHIFT;
				const int64_t inv_z_max = 1 << 29;
				const uint32_t span_inv_z_start =
					uint32_t(std::min(std::max(span_start_x * d_inv_z_dx + line_inv_z, inv_z_min), inv_z_max));
				const uint32_t span_inv_z_end =
					uint32_t(std::min(std::max(span_end_x * d_inv_z_dx + line_inv_z, inv_z_min), inv_z_max));
				const uint64_t span_z_start = (1u << Z_CALC_SHIFT) / (span_inv_z_start >> INV_Z_PRE_SHIFT);
				const uint64_t span_z_end = (1u << Z_CALC_SHIFT) / (span_inv_z_end >> INV_Z_PRE_SHIFT);

				int32_t span_tc[2] = {0, 0}, span_d_tc[2] = {0, 0};
				for(int i = 0; i < 2; ++i)
				{
					const uint64_t tc_start = uint64_t(std::max<int64_t>(line_tc[i] + span_start_x * d_tc_dx[i], 0));
					const uint64_t tc_end = uint64_t(std::max<int64_t>(line_tc[i] + span_end_x * d_tc_dx[i], 0));
					const uint64_t max_tc = (uint64_t(uint32_t(tsize[i])) << LINE_TC_SHIFT) - 1;
					const int32_t tc_mul_z_start =
						int32_t(std::min((span_z_start * tc_start) >> (TC_FINAL_SHIFT - LINE_TC_SHIFT), max_tc));
					const int32_t tc_mul_z_end =
						int32_t(std::min((span_z_end * tc_end) >> (TC_FINAL_SHIFT - LINE_TC_SHIFT), max_tc));
					span_tc[i] = tc_mul_z_start;
					if(span_length_minus_one > 0)
						span_d_tc[i] = (tc_mul_z_end - tc_mul_z_start) / span_length_minus_one;
				}

				for(int32_t x = x_start_int; x < x_end_int; ++x)
				{
					const int32_t tc0 = span_tc[0] >> LINE_TC_SHIFT;
					const int32_t tc1 = span_tc[1] >> LINE_TC_SHIFT;
					if(in_texture(uint32_t(tc0), uint32_t(tc1), tsize))
						write_into_framebuffer<CT>(
							blend, line_dst[x], texture_data[uint32_t(tc0) + uint32_t(tc1) * texture_width]);
					span_tc[0] += span_d_tc[0];
					span_tc[1] += span_d_tc[1];
				}
				count(blend, x_end_int - x_start_int, false);
			}
			x_left += left_side.dx_dy;
			x_right += right_side.dx_dy;
			line_inv_z += d_inv_z_dy;
			line_tc[0] += d_tc_dy[0];
			line_tc[1] += d_tc_dy[1];
		}
	}

	// rasterizer.rs:583-729
	void fill_polygon_part_affine(
		uint32_t blend, Fixed16 y_start, Fixed16 y_end, PolygonSide left_side, PolygonSide right_side, const float* de,
		const float* te, const int32_t* tsize, const typename CT::T* texture_data)
	{
		const float y_start_f32 = fixed16_to_f32(y_start);
		const float y_end_f32 = fixed16_to_f32(y_end);
		const Fixed16 y_delta = y_end - y_start;
		const float x_start_left = fixed16_to_f32(left_side.x_start);
		const float x_end_left = fixed16_to_f32(left_side.x_start + fixed16_mul(y_delta, left_side.dx_dy));
		const float x_start_right = fixed16_to_f32(right_side.x_start);
		const float x_end_right = fixed16_to_f32(right_side.x_start + fixed16_mul(y_delta, right_side.dx_dy));
		const float z_start_left = 1.0f / (de[0] * x_start_left + de[1] * y_start_f32 + de[2]);
		const float z_end_left = 1.0f / (de[0] * x_end_left + de[1] * y_end_f32 + de[2]);
		const float z_start_right = 1.0f / (de[0] * x_start_right + de[1] * y_start_f32 + de[2]);
		const float z_end_right = 1.0f / (de[0] * x_end_right + de[1] * y_end_f32 + de[2]);
		const Fixed16 y_delta_for_tc_interpolation = std::max(y_delta, FIXED16_ONE);

		Fixed16 tc_left[2], tc_right[2], d_tc_left[2], d_tc_right[2];
		for(int i = 0; i < 2; ++i)
		{
			const Fixed16 max_tc = int_to_fixed16(tsize[i]) - 1;
			auto conv = [&](float z, float x, float y) {
				return std::min(
					std::max(f32_to_fixed16_unchecked(z * (x * te[0 + i] + y * te[2 + i] + te[4 + i])), 0), max_tc);
			};
			const Fixed16 tc_start_left = conv(z_start_left, x_start_left, y_start_f32);
			const Fixed16 tc_end_left = conv(z_end_left, x_end_left, y_end_f32);
			const Fixed16 tc_start_right = conv(z_start_right, x_start_right, y_start_f32);
			const Fixed16 tc_end_right = conv(z_end_right, x_end_right, y_end_f32);
			d_tc_left[i] = fixed16_div(tc_end_left - tc_start_left, y_delta_for_tc_interpolation);
			d_tc_right[i] = fixed16_div(tc_end_right - tc_start_right, y_delta_for_tc_interpolation);
			tc_left[i] = tc_start_left;
			tc_right[i] = tc_start_right;
		}

		const int32_t y_start_int = std::max(fixed16_round_to_int(y_start), clip_rect.min_y);
		const int32_t y_end_int = std::min(fixed16_round_to_int(y_end), clip_rect.max_y);
		const Fixed16 y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y_start;
		Fixed16 x_left = left_side.x_start + fixed16_mul(y_start_delta, left_side.dx_dy) + FIXED16_HALF;
		Fixed16 x_right = right_side.x_start + fixed16_mul(y_start_delta, right_side.dx_dy) + FIXED16_HALF;
		for(int i = 0; i < 2; ++i)
		{
			tc_left[i] += fixed16_mul(y_start_delta, d_tc_left[i]);
			tc_right[i] += fixed16_mul(y_start_delta, d_tc_right[i]);
		}

		for(int32_t y_int = y_start_int; y_int < y_end_int; ++y_int)
		{
			const int32_t x_start_int = std::max(fixed16_floor_to_int(x_left), clip_rect.min_x);
			const int32_t x_end_int = std::min(fixed16_floor_to_int(x_right), clip_rect.max_x);
			if(x_start_int < x_end_int)
			{
				typename CT::T* line_dst = color_buffer + int64_t(y_int) * row_size;
				Fixed16 tc[2] = {tc_left[0], tc_left[1]};
				const Fixed16 x_delta = std::max(x_right - x_left, FIXED16_ONE);
				const Fixed16 d_tc[2] = {
					fixed16_div(tc_right[0] - tc_left[0], x_delta), fixed16_div(tc_right[1] - tc_left[1], x_delta)};
				for(int32_t x = x_start_int; x < x_end_int; ++x)
				{
					const int32_t u = fixed16_floor_to_int(tc[0]);
					const int32_t v = fixed16_floor_to_int(tc[1]);
					if(in_texture(uint32_t(u), uint32_t(v), tsize))
						write_into_framebuffer<CT>(blend, line_dst[x], texture_data[u + v * tsize[0]]);
					tc[0] += d_tc[0];
					tc[1] += d_tc[1];
				}
				count(blend, x_end_int - x_start_int, false);
			}
			x_left += left_side.dx_dy;
			x_right += right_side.dx_dy;
			tc_left[0] += d_tc_left[0];
			tc_left[1] += d_tc_left[1];
			tc_right[0] += d_tc_right[0];
			tc_right[1] += d_tc_right[1];
		}
	}

	void draw_part(
		uint32_t mode, uint32_t blend, Fixed16 y0, Fixed16 y1, PolygonSide l, PolygonSide r, const float* de,
		const float* te, const int32_t* ts, const typename CT::T* tex)
	{
		if(mode == SW_TC_LINE_Z)
			fill_polygon_part_line_z_corrected(blend, y0, y1, l, r, de, te, ts, tex);
		else if(mode == SW_TC_AFFINE)
			fill_polygon_part_affine(blend, y0, y1, l, r, de, te, ts, tex);
		else
			fill_polygon_part(blend, y0, y1, l, r, de, te, ts, tex);
	}

	// rasterizer.rs:137-271
	void fill_polygon(
		const int32_t* vertices, uint32_t n, uint32_t mode, uint32_t blend, const float* de, const float* te,
		const int32_t* tsize, const typename CT::T* texture_data)
	{
		auto vx = [&](uint32_t i) { return vertices[i * 2 + 0]; };
		auto vy = [&](uint32_t i) { return vertices[i * 2 + 1]; };
		uint32_t lower_vertex_index = 0;
		Fixed16 min_y = vy(0);
		for(uint32_t i = 0; i < n; ++i)
			if(vy(i) < min_y)
			{
				min_y = vy(i);
				lower_vertex_index = i;
			}

		uint32_t left_index = lower_vertex_index, right_index = lower_vertex_index;
		Fixed16 cur_y = min_y;
		for(;;)
		{
			uint32_t next_left_index = left_index + n - 1;
			if(next_left_index >= n)
				next_left_index -= n;
			uint32_t next_right_index = right_index + 1;
			if(next_right_index >= n)
				next_right_index -= n;

			const Fixed16 dy_left = vy(next_left_index) - vy(left_index);
			const Fixed16 dy_right = vy(next_right_index) - vy(right_index);
			const Fixed16 next_y = std::min(vy(next_left_index), vy(next_right_index));
			if(dy_left > FIXED16_HALF && dy_right > FIXED16_HALF)
			{
				const Fixed16 dx_dy_left = fixed16_div(vx(next_left_index) - vx(left_index), dy_left);
				const Fixed16 dx_dy_right = fixed16_div(vx(next_right_index) - vx(right_index), dy_right);
				draw_part(
					mode, blend, cur_y, next_y,
					PolygonSide{vx(left_index) + fixed16_mul(dx_dy_left, cur_y - vy(left_index)), dx_dy_left},
					PolygonSide{vx(right_index) + fixed16_mul(dx_dy_right, cur_y - vy(right_index)), dx_dy_right}, de,
					te, tsize, texture_data);
			}
			else if(dy_left > 0 && dy_right > 0)
			{
				const int32_t cur_y_int = fixed16_round_to_int(cur_y);
				const int32_t next_y_int = fixed16_round_to_int(next_y);
				if(cur_y_int < next_y_int)
				{
					const Fixed16 thin_line_y = int_to_fixed16(cur_y_int) + FIXED16_HALF;
					const Fixed16 x_start_left = vx(left_index) +
						fixed16_mul_div(thin_line_y - vy(left_index), vx(next_left_index) - vx(left_index), dy_left);
					const Fixed16 x_start_right = vx(right_index) +
						fixed16_mul_div(thin_line_y - vy(right_index), vx(next_right_index) - vx(right_index), dy_right);
					draw_part(
						mode, blend, cur_y, next_y, PolygonSide{x_start_left, 0}, PolygonSide{x_start_right, 0}, de, te,
						tsize, texture_data);
				}
			}

			if(next_left_index == next_right_index)
				break;
			if(vy(next_right_index) < vy(next_left_index))
				right_index = next_right_index;
			else
				left_index = next_left_index;
			cur_y = next_y;
		}
	}

	// rasterizer.rs:984-1292
	void fill_triangle_part(
		uint32_t blend, Fixed16 y_start, Fixed16 y_end, PolygonSide left_side, PolygonSide right_side,
		const Fixed16 tc_start_left[2], Fixed16 d_tc_dy_left[2], Fixed16 d_tc_dx[2], const Fixed16 light_start_left[3],
		Fixed16 d_light_left[3], Fixed16 d_light_dx[3], const int32_t* tsize, const uint32_t* texture_data)
	{
		const int32_t y_start_int = std::max(fixed16_round_to_int(y_start), clip_rect.min_y);
		const int32_t y_end_int = std::min(fixed16_round_to_int(y_end), clip_rect.max_y);
		if(y_start_int >= y_end_int)
			return;

		const Fixed16 y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y_start;
		Fixed16 x_left = left_side.x_start + fixed16_mul(y_start_delta, left_side.dx_dy) + FIXED16_HALF;
		Fixed16 x_right = right_side.x_start + fixed16_mul(y_start_delta, right_side.dx_dy) + FIXED16_HALF;

		const Fixed16 y_end_delta = y_end - (y_start + y_start_delta);
		const Fixed16 d_x_lower = std::max(x_right - x_left, 0);
		const Fixed16 x_upper_left = x_left + fixed16_mul(y_end_delta, left_side.dx_dy);
		const Fixed16 x_upper_right = x_right + fixed16_mul(y_end_delta, right_side.dx_dy);
		const Fixed16 d_x_upper = std::max(x_upper_right - x_upper_left, 0);

		Fixed16 tc_left[2] = {0, 0};
		for(int i = 0; i < 2; ++i)
		{
			const Fixed16 max_tc = int_to_fixed16(tsize[i]) - 1;
			{
				tc_left[i] = std::min(std::max(tc_start_left[i] + fixed16_mul(y_start_delta, d_tc_dy_left[i]), 0), max_tc);
				const Fixed16 tc_lower_delta = fixed16_mul(d_tc_dx[i], d_x_lower);
				if(tc_lower_delta > max_tc)
				{
					tc_left[i] = 0;
					d_tc_dx[i] = fixed16_div(max_tc, d_x_lower);
				}
				else if(tc_lower_delta < -max_tc)
				{
					tc_left[i] = max_tc;
					d_tc_dx[i] = fixed16_div(-max_tc, d_x_lower);
				}
				else
				{
					const Fixed16 tc_lower_right = tc_left[i] + tc_lower_delta;
					if(tc_lower_right > max_tc)
						tc_left[i] -= tc_lower_right - max_tc;
					else if(tc_lower_right < 0)
						tc_left[i] -= tc_lower_right;
				}
			}
			{
				Fixed16 tc_upper_left =
					std::min(std::max(tc_left[i] + fixed16_mul(y_end_delta, d_tc_dy_left[i]), 0), max_tc);
				const Fixed16 tc_upper_delta = fixed16_mul(d_tc_dx[i], d_x_upper);
				if(tc_upper_delta > max_tc)
				{
					tc_upper_left = 0;
					d_tc_dx[i] = fixed16_div(max_tc, d_x_upper);
				}
				else if(tc_upper_delta < -max_tc)
				{
					tc_upper_left = max_tc;
					d_tc_dx[i] = fixed16_div(-max_tc, d_x_upper);
				}
				else
				{
					const Fixed16 tc_upper_right = tc_upper_left + tc_upper_delta;
					if(tc_upper_right > max_tc)
						tc_upper_left -= tc_upper_right - max_tc;
					else if(tc_upper_right < 0)
						tc_upper_left -= tc_upper_right;
				}
				if(y_end_delta > 0)
					d_tc_dy_left[i] = fixed16_div(tc_upper_left - tc_left[i], y_end_delta);
				else
					d_tc_dy_left[i] = 0;

				const Fixed16 tc_upper_left2 = tc_left[i] + fixed16_mul(y_end_delta, d_tc_dy_left[i]);
				const Fixed16 tc_upper_delta2 = fixed16_mul(d_tc_dx[i], d_x_upper);
				const Fixed16 tc_upper_right2 = tc_upper_left2 + tc_upper_delta2;
				if(tc_upper_right2 < 0)
					d_tc_dx[i] = d_x_upper > 0 ? fixed16_div(-tc_upper_left2, d_x_upper) : 0;
				else if(tc_upper_right2 > max_tc)
					d_tc_dx[i] = d_x_upper > 0 ? fixed16_div(max_tc - tc_upper_left2, d_x_upper) : 0;
			}
		}

		Fixed16 light_left[3] = {0, 0, 0};
		for(int i = 0; i < 3; ++i)
		{
			light_left[i] = std::max(light_start_left[i] + fixed16_mul(y_start_delta, d_light_left[i]), 0);
			const Fixed16 light_lower_right = light_left[i] + fixed16_mul(d_x_lower, d_light_dx[i]);
			if(light_lower_right < 0)
				light_left[i] -= light_lower_right;
			{
				Fixed16 light_upper_left = std::max(light_left[i] + fixed16_mul(y_end_delta, d_light_left[i]), 0);
				const Fixed16 light_upper_right = light_upper_left + fixed16_mul(d_x_upper, d_light_dx[i]);
				if(light_upper_right < 0)
					light_upper_left -= light_upper_right;
				if(y_end_delta > 0)
					d_light_left[i] = fixed16_div(light_upper_left - light_left[i], y_end_delta);
				else
					d_light_left[i] = 0;
				const Fixed16 light_upper_left2 = light_left[i] + fixed16_mul(y_end_delta, d_light_left[i]);
				const Fixed16 light_upper_right2 = light_upper_left2 + fixed16_mul(d_x_upper, d_light_dx[i]);
				if(light_upper_right2 < 0)
					d_light_dx[i] = d_x_upper > 0 ? fixed16_div(-light_upper_left2, d_x_upper) : 0;
			}
		}

		const __m128i d_light_dx_vec = ci_from_i32x3(d_light_dx);
		for(int32_t y_int = y_start_int; y_int < y_end_int; ++y_int)
		{
			const int32_t x_start_int = std::max(fixed16_floor_to_int(x_left), clip_rect.min_x);
			const int32_t x_end_int = std::min(fixed16_floor_to_int(x_right), clip_rect.max_x);
			if(x_start_int < x_end_int)
			{
				const Fixed16 x_start_delta = int_to_fixed16(x_start_int) + FIXED16_ONE - x_left;
				Fixed16 line_tc[2] = {
					tc_left[0] + fixed16_mul(x_start_delta, d_tc_dx[0]), tc_left[1] + fixed16_mul(x_start_delta, d_tc_dx[1])};
				const int32_t ll[3] = {
					light_left[0] + fixed16_mul(x_start_delta, d_light_dx[0]),
					light_left[1] + fixed16_mul(x_start_delta, d_light_dx[1]),
					light_left[2] + fixed16_mul(x_start_delta, d_light_dx[2])};
				__m128i line_light = ci_from_i32x3(ll);
				typename CT::T* line_dst = color_buffer + int64_t(y_int) * row_size;
				for(int32_t x = x_start_int; x < x_end_int; ++x)
				{
					const int32_t u = fixed16_floor_to_int(line_tc[0]);
					const int32_t v = fixed16_floor_to_int(line_tc[1]);
					if(!in_texture(uint32_t(u), uint32_t(v), tsize))
					{
						line_tc[0] += d_tc_dx[0];
						line_tc[1] += d_tc_dx[1];
						line_light = _mm_add_epi32(line_light, d_light_dx_vec);
						continue;
					}
					const uint32_t texel = texture_data[u + v * tsize[0]];
					const __m128i lighted = _mm_srai_epi32(_mm_mullo_epi32(ci_from_color32(texel), line_light), 16);
					typename CT::T& dst = line_dst[x];
					switch(blend)
					{
					case SW_BLEND_NONE: dst = CT::from_veci(lighted); break;
					case SW_BLEND_AVERAGE: dst = CT::average(dst, CT::from_veci(lighted)); break;
					case SW_BLEND_ADDITIVE: dst = CT::saturated_sum(dst, CT::from_veci(lighted)); break;
					case SW_BLEND_ALPHA_TEST:
						if(C32::test_alpha(texel))
							dst = CT::from_veci(lighted);
						break;
					case SW_BLEND_ALPHA_BLEND:
					{
						const int32_t alpha = C32::get_alpha(texel);
						dst = CT::from_veci(_mm_srai_epi32(
							_mm_add_epi32(
								_mm_mullo_epi32(CT::to_veci(dst), _mm_set1_epi32(255 - alpha)),
								_mm_mullo_epi32(lighted, _mm_set1_epi32(alpha))),
							8));
						break;
					}
					default: break;
					}
					line_tc[0] += d_tc_dx[0];
					line_tc[1] += d_tc_dx[1];
					line_light = _mm_add_epi32(line_light, d_light_dx_vec);
				}
				count(blend, x_end_int - x_start_int, true);
			}
			x_left += left_side.dx_dy;
			x_right += right_side.dx_dy;
			tc_left[0] += d_tc_dy_left[0];
			tc_left[1] += d_tc_dy_left[1];
			light_left[0] += d_light_left[0];
			light_left[1] += d_light_left[1];
			light_left[2] += d_light_left[2];
		}
	}

	// rasterizer.rs:765-982
	void fill_triangle(const int32_t* vdata, uint32_t blend, const int32_t* tsize, const uint32_t* texture_data)
	{
		struct TV
		{
			Fixed16 x, y, tc[2], light[3];
		};
		TV v[3];
		for(int i = 0; i < 3; ++i)
		{
			const int32_t* p = vdata + i * 7;
			v[i] = TV{p[0], p[1], {p[2], p[3]}, {p[4], p[5], p[6]}};
		}
		int upper_index, lower_index;
		if(v[0].y >= v[1].y && v[0].y >= v[2].y)
		{
			upper_index = 0;
			lower_index = v[1].y < v[2].y ? 1 : 2;
		}
		else if(v[1].y >= v[0].y && v[1].y >= v[2].y)
		{
			upper_index = 1;
			lower_index = v[0].y < v[2].y ? 0 : 2;
		}
		else
		{
			upper_index = 2;
			lower_index = v[0].y < v[1].y ? 0 : 1;
		}
		const int middle_index = 3 - upper_index - lower_index;
		const TV& up = v[upper_index];
		const TV& lo = v[lower_index];
		const TV& mi = v[middle_index];

		const Fixed16 long_edge_dy = std::max(up.y - lo.y, FIXED16_HALF);
		const Fixed16 lower_part_dy = std::max(mi.y - lo.y, FIXED16_HALF);
		const Fixed16 upper_part_dy = std::max(up.y - mi.y, FIXED16_HALF);
		const Fixed16 long_edge_dx_dy = fixed16_div(up.x - lo.x, long_edge_dy);
		const Fixed16 long_edge_x_in_middle = lo.x + fixed16_mul(long_edge_dx_dy, mi.y - lo.y);
		const Fixed16 lower_part_dx_dy = fixed16_div(mi.x - lo.x, lower_part_dy);
		const Fixed16 upper_part_dx_dy = fixed16_div(up.x - mi.x, upper_part_dy);

		Fixed16 long_edge_d_tc_dy[2], long_edge_tc_in_middle[2], d_tc_dy_lower[2], d_tc_dy_upper[2];
		for(int i = 0; i < 2; ++i)
		{
			long_edge_d_tc_dy[i] = fixed16_div(up.tc[i] - lo.tc[i], long_edge_dy);
			long_edge_tc_in_middle[i] = lo.tc[i] + fixed16_mul(long_edge_d_tc_dy[i], lower_part_dy);
			d_tc_dy_lower[i] = fixed16_div(mi.tc[i] - lo.tc[i], lower_part_dy);
			d_tc_dy_upper[i] = fixed16_div(up.tc[i] - mi.tc[i], upper_part_dy);
		}
		Fixed16 long_edge_d_light_dy[3], long_edge_light_in_middle[3], d_light_dy_lower[3], d_light_dy_upper[3];
		for(int i = 0; i < 3; ++i)
		{
			long_edge_d_light_dy[i] = fixed16_div(up.light[i] - lo.light[i], long_edge_dy);
			long_edge_light_in_middle[i] = lo.light[i] + fixed16_mul(long_edge_d_light_dy[i], lower_part_dy);
			d_light_dy_lower[i] = fixed16_div(mi.light[i] - lo.light[i], lower_part_dy);
			d_light_dy_upper[i] = fixed16_div(up.light[i] - mi.light[i], upper_part_dy);
		}

		if(long_edge_x_in_middle >= v[middle_index].x)
		{
			const Fixed16 middle_dx = std::max(long_edge_x_in_middle - mi.x, FIXED16_HALF);
			Fixed16 d_tc_dx[2], d_light_dx[3];
			for(int i = 0; i < 2; ++i)
				d_tc_dx[i] = fixed16_div(long_edge_tc_in_middle[i] - mi.tc[i], middle_dx);
			for(int i = 0; i < 3; ++i)
				d_light_dx[i] = fixed16_div(long_edge_light_in_middle[i] - mi.light[i], middle_dx);
			{
				Fixed16 dtcy[2] = {d_tc_dy_lower[0], d_tc_dy_lower[1]}, dtcx[2] = {d_tc_dx[0], d_tc_dx[1]};
				Fixed16 dly[3] = {d_light_dy_lower[0], d_light_dy_lower[1], d_light_dy_lower[2]};
				Fixed16 dlx[3] = {d_light_dx[0], d_light_dx[1], d_light_dx[2]};
				fill_triangle_part(
					blend, lo.y, mi.y, PolygonSide{lo.x, lower_part_dx_dy}, PolygonSide{lo.x, long_edge_dx_dy}, lo.tc, dtcy,
					dtcx, lo.light, dly, dlx, tsize, texture_data);
			}
			{
				Fixed16 dtcy[2] = {d_tc_dy_upper[0], d_tc_dy_upper[1]}, dtcx[2] = {d_tc_dx[0], d_tc_dx[1]};
				Fixed16 dly[3] = {d_light_dy_upper[0], d_light_dy_upper[1], d_light_dy_upper[2]};
				Fixed16 dlx[3] = {d_light_dx[0], d_light_dx[1], d_light_dx[2]};
				fill_triangle_part(
					blend, mi.y, up.y, PolygonSide{mi.x, upper_part_dx_dy}, PolygonSide{long_edge_x_in_middle, long_edge_dx_dy},
					mi.tc, dtcy, dtcx, mi.light, dly, dlx, tsize, texture_data);
			}
		}
		else
		{
			const Fixed16 middle_dx = std::max(mi.x - long_edge_x_in_middle, FIXED16_HALF);
			Fixed16 d_tc_dx[2], d_light_dx[3];
			for(int i = 0; i < 2; ++i)
				d_tc_dx[i] = fixed16_div(mi.tc[i] - long_edge_tc_in_middle[i], middle_dx);
			for(int i = 0; i < 3; ++i)
				d_light_dx[i] = fixed16_div(mi.light[i] - long_edge_light_in_middle[i], middle_dx);
			{
				Fixed16 dtcy[2] = {long_edge_d_tc_dy[0], long_edge_d_tc_dy[1]}, dtcx[2] = {d_tc_dx[0], d_tc_dx[1]};
				Fixed16 dly[3] = {long_edge_d_light_dy[0], long_edge_d_light_dy[1], long_edge_d_light_dy[2]};
				Fixed16 dlx[3] = {d_light_dx[0], d_light_dx[1], d_light_dx[2]};
				fill_triangle_part(
					blend, lo.y, mi.y, PolygonSide{lo.x, long_edge_dx_dy}, PolygonSide{lo.x, lower_part_dx_dy}, lo.tc, dtcy,
					dtcx, lo.light, dly, dlx, tsize, texture_data);
			}
			{
				Fixed16 dtcy[2] = {long_edge_d_tc_dy[0], long_edge_d_tc_dy[1]}, dtcx[2] = {d_tc_dx[0], d_tc_dx[1]};
				Fixed16 dly[3] = {long_edge_d_light_dy[0], long_edge_d_light_dy[1], long_edge_d_light_dy[2]};
				Fixed16 dlx[3] = {d_light_dx[0], d_light_dx[1], d_light_dx[2]};
				fill_triangle_part(
					blend, mi.y, up.y, PolygonSide{long_edge_x_in_middle, long_edge_dx_dy}, PolygonSide{mi.x, upper_part_dx_dy},
					long_edge_tc_in_middle, dtcy, dtcx, long_edge_light_in_middle, dly, dlx, tsize, texture_data);
			}
		}
	}
};

// ---------------------------------------------------------------------------------------------
// Frame driver over a batch of views
// ---------------------------------------------------------------------------------------------
template <class CT>
void build_view_surfaces(
	const Numerics& nm, const MapData& m, const sw_batch& b, const sw_view& view, typename CT::T* arena,
	uint32_t first, uint32_t end, Counters* counters)
{
	const Vec3 cam = v3(view.cam_pos[0], view.cam_pos[1], view.cam_pos[2]);
	for(uint32_t si = first; si < end; ++si)
	{
		const sw_surface_desc& s = b.surfaces[si];
		typename CT::T* out = arena + view.surface_arena_offset + s.pixels_offset;
		if(s.flags & SW_SURF_DIFFUSE_LAYER)
			build_surface<CT>(nm, m, s, b.lights, b.light_indices, cam, out, counters);
		if((s.flags & SW_SURF_EMISSIVE) && s.emissive_level != SW_NO_TEXTURE)
			mix_surface_with_texture<CT>(m, s, s.emissive_blend, out, counters);
	}
}

template <class CT>
void raster_view(
	const MapData& m, const sw_batch& b, const sw_view& view, const typename CT::T* arena, typename CT::T* fb,
	const ClipRect* override_clip, Counters* counters)
{
	Rasterizer<CT> r;
	r.color_buffer = fb;
	r.row_size = int32_t(view.width);
	r.counters = counters;
	for(uint32_t di = view.first_draw; di < view.first_draw + view.num_draws; ++di)
	{
		const sw_draw& d = b.draws[di];
		r.clip_rect = override_clip ? *override_clip : ClipRect{d.clip_rect[0], d.clip_rect[1], d.clip_rect[2], d.clip_rect[3]};
		const int32_t* vd = b.vertex_data + d.first_vertex_int;
		if(d.kind == SW_DRAW_POLYGON)
		{
			if(d.num_vertices < 3)
				continue;
			const typename CT::T* tex = d.texture_source == SW_TEXSRC_SKYBOX ?
				static_cast<const typename CT::T*>(m.skybox) + d.texture_offset :
				d.texture_source == SW_TEXSRC_PORTAL ? static_cast<const typename CT::T*>(b.portal_texels) + d.texture_offset :
				arena + view.surface_arena_offset + d.texture_offset;
			if(d.texture_source == SW_TEXSRC_PORTAL &&
				(!b.portal_texels ||
					d.texture_offset + uint64_t(uint32_t(d.texture_size[0])) * uint32_t(d.texture_size[1]) > b.num_portal_texels))
				continue; // outside the portal textures: skipped, as the kernels skip it
			if(d.texture_source > SW_TEXSRC_PORTAL)
				continue;
			r.fill_polygon(vd, d.num_vertices, d.mode, d.blend, d.depth_equation, d.tc_equation, d.texture_size, tex);
		}
		else
			r.fill_triangle(vd, d.blend, d.texture_size, m.lite_texels + d.texture_offset);
	}
}

// rect_splitting.rs:29-78
struct RectF
{
	float min_x, min_y, max_x, max_y;
};
void split_rect_r(const RectF& r, uint32_t num_parts, std::vector<RectF>& out)
{
	if(num_parts <= 1)
	{
		out.push_back(r);
		return;
	}
	const uint32_t n0 = num_parts / 2, n1 = num_parts - n0;
	const float ratio0 = float(n0) / float(num_parts), ratio1 = float(n1) / float(num_parts);
	RectF r0, r1;
	if(r.max_x - r.min_x >= r.max_y - r.min_y)
	{
		const float middle = r.max_x * ratio0 + r.min_x * ratio1;
		r0 = RectF{r.min_x, r.min_y, middle, r.max_y};
		r1 = RectF{middle, r.min_y, r.max_x, r.max_y};
	}
	else
	{
		const float middle = r.max_y * ratio0 + r.min_y * ratio1;
		r0 = RectF{r.min_x, r.min_y, r.max_x, middle};
		r1 = RectF{r.min_x, middle, r.max_x, r.max_y};
	}
	split_rect_r(r0, n0, out);
	split_rect_r(r1, n1, out);
}

template <class CT>
int render_batch(
	int mode, const sw_map_desc* map, const sw_batch* b, void* arena_v, void* fb_v, uint32_t num_threads,
	sw_stats* stats, uint64_t* out_of_bounds)
{
	Numerics nm{mode};
	MapData m;
	m.texels = map->texels;
	m.lite_texels = map->lite_texels;
	m.lightmaps_simple = map->lightmaps_simple;
	m.lightmaps_directional = map->lightmaps_directional;
	m.levels = map->levels;
	m.skybox = sizeof(typename CT::T) == 4 ? static_cast<const void*>(map->skybox_texels32) :
											static_cast<const void*>(map->skybox_texels64);
	typename CT::T* arena = static_cast<typename CT::T*>(arena_v);
	typename CT::T* fb = static_cast<typename CT::T*>(fb_v);
	Counters counters;
	Counters* cptr = &counters;

	for(uint32_t vi = 0; vi < b->num_views; ++vi)
	{
		const sw_view& view = b->views[vi];
		typename CT::T* vfb = fb + view.fb_offset;
		if(num_threads <= 1)
		{
			build_view_surfaces<CT>(nm, m, *b, view, arena, view.first_surface, view.first_surface + view.num_surfaces, cptr);
			if(view.clear_background)
				std::fill(vfb, vfb + uint64_t(view.width) * view.height, typename CT::T(view.background_color));
			raster_view<CT>(m, *b, view, arena, vfb, nullptr, cptr);
		}
		else
		{
			// Reference decomposition: surfaces in parallel (partial_renderer.rs:1822-1831), screen split
			// into num_threads rects each rasterised with its own clip rect (:1235-1282).
			std::atomic<uint32_t> next{view.first_surface};
			const uint32_t end = view.first_surface + view.num_surfaces;
			std::vector<std::thread> pool;
			std::vector<Counters> local(std::max<uint32_t>(num_threads, 64));
			for(uint32_t t = 0; t < num_threads; ++t)
				pool.emplace_back([&, t]() {
					for(;;)
					{
						const uint32_t i = next.fetch_add(4);
						if(i >= end)
							break;
						build_view_surfaces<CT>(nm, m, *b, view, arena, i, std::min(i + 4, end), &local[t]);
					}
				});
			for(auto& t : pool)
				t.join();
			pool.clear();
			if(view.clear_background)
				std::fill(vfb, vfb + uint64_t(view.width) * view.height, typename CT::T(view.background_color));
			std::vector<RectF> rects;
			split_rect_r(RectF{0.0f, 0.0f, float(view.width), float(view.height)}, std::min<uint32_t>(num_threads, 64), rects);
			std::vector<Counters> rect_local(rects.size());
			for(size_t k = 0; k < rects.size(); ++k)
				pool.emplace_back([&, k]() {
					const RectF& rf = rects[k];
					const ClipRect c{int32_t(rf.min_x), int32_t(rf.min_y), int32_t(rf.max_x), int32_t(rf.max_y)};
					raster_view<CT>(m, *b, view, arena, vfb, &c, &rect_local[k]);
				});
			for(auto& t : pool)
				t.join();
			for(const Counters& c : local)
				counters.add(c);
			for(const Counters& c : rect_local)
				counters.add(c);
		}
	}
	if(stats)
	{
		stats->surface_texels = counters.surface_texels;
		stats->lightmap_elements = counters.lightmap_elements;
		stats->emissive_texels = counters.emissive_texels;
		stats->polygon_pixels = counters.polygon_pixels;
		stats->triangle_pixels = counters.triangle_pixels;
		stats->blend_pixels = counters.blend_pixels;
	}
	if(out_of_bounds)
		*out_of_bounds = counters.out_of_bounds;
	return 0;
}

// postprocessor.rs:226-256 + :1003-1024
void tonemap_rows(
	const uint64_t* hdr, uint32_t width, uint32_t y0, uint32_t y1, float exposure, const float modulate[3], uint32_t* dst,
	uint32_t dst_pitch, __m128* sum_out)
{
	const float inv_255 = 1.0f / 255.0f;
	const float inv_scale[3] = {1.0f / (exposure * modulate[0]), 1.0f / (exposure * modulate[1]), 1.0f / (exposure * modulate[2])};
	const float i255[3] = {inv_255, inv_255, inv_255};
	const __m128 inv_scale_vec = cv_from_f32x3(inv_scale);
	const __m128 inv_255_vec = cv_from_f32x3(i255);
	__m128 colors_sum = _mm_setzero_ps();
	for(uint32_t y = y0; y < y1; ++y)
	{
		__m128 line_sum = _mm_setzero_ps();
		for(uint32_t x = 0; x < width; ++x)
		{
			const __m128 c = cv_from_color64(hdr[uint64_t(y) * width + x]);
			line_sum = _mm_add_ps(line_sum, c);
			const __m128 mapped = _mm_div_ps(c, _mm_fmadd_ps(c, inv_255_vec, inv_scale_vec));
			dst[uint64_t(y) * dst_pitch + x] = cv_into_color32(mapped);
		}
		colors_sum = _mm_add_ps(colors_sum, line_sum);
	}
	*sum_out = colors_sum;
}

// ---------------------------------------------------------------------------------------------
// Bloom postprocessing (postprocessor.rs:258-956), restated loop by loop, single-threaded (the
// reference's row ranges partition the same per-pixel work).
// ---------------------------------------------------------------------------------------------
const int MAX_BLOOM_RADIUS = 16; // postprocessor.rs:941
const int MAX_BLOOM_KERNEL = 1 + 2 * MAX_BLOOM_RADIUS;

// compute_gaussian_kernel (postprocessor.rs:1026-1045)
void gaussian_kernel(float sigma, int radius, float out[MAX_BLOOM_KERNEL])
{
	for(int i = 0; i < MAX_BLOOM_KERNEL; ++i)
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

struct Bloom
{
	uint32_t w = 0, h = 0; // hdr size
	uint32_t bw = 0, bh = 0, scale = 1;
	std::vector<uint64_t> buf[2];
	__m128 average = _mm_setzero_ps();
};

// downscale_hdr_buffer (postprocessor.rs:805-853): integer box average into buffer 0, and the
// average colour (integer line sums, float sum of lines)
void bloom_downscale(const uint64_t* hdr, Bloom& b)
{
	const int color_shift = 8;
	const int32_t average_scaler = (1 << color_shift) / int32_t(b.scale * b.scale);
	__m128 colors_sum = _mm_setzero_ps();
	for(uint32_t dy = 0; dy < b.bh; ++dy)
	{
		__m128i line_sum = _mm_setzero_si128();
		for(uint32_t dx = 0; dx < b.bw; ++dx)
		{
			__m128i sum = _mm_setzero_si128();
			for(uint32_t y = 0; y < b.scale; ++y)
				for(uint32_t x = 0; x < b.scale; ++x)
					sum = _mm_add_epi32(sum, ci_from_color64(hdr[uint64_t(dy * b.scale + y) * b.w + dx * b.scale + x]));
			const __m128i average = _mm_srai_epi32(_mm_mullo_epi32(sum, _mm_set1_epi32(average_scaler)), color_shift);
			b.buf[0][uint64_t(dy) * b.bw + dx] = ci_into_color64(average);
			line_sum = _mm_add_epi32(line_sum, average);
		}
		colors_sum = _mm_add_ps(colors_sum, _mm_cvtepi32_ps(line_sum));
	}
	b.average = _mm_mul_ps(colors_sum, _mm_set1_ps(1.0f / float(b.bw * b.bh)));
}

// perform_blur_impl (postprocessor.rs:855-937): horizontal pass 0 -> 1, vertical pass 1 -> 0 with
// the weights scaled by bloom_scale
void bloom_blur(Bloom& b, const float kernel[MAX_BLOOM_KERNEL], int radius, float bloom_scale)
{
	const int color_shift = 14;
	int32_t ki[MAX_BLOOM_KERNEL];
	for(int i = 0; i < MAX_BLOOM_KERNEL; ++i)
		ki[i] = int32_t(kernel[i] * float(1 << color_shift));
	const int32_t last_x = int32_t(b.bw) - 1, last_y = int32_t(b.bh) - 1;
	for(uint32_t y = 0; y < b.bh; ++y)
		for(uint32_t x = 0; x < b.bw; ++x)
		{
			const uint64_t* line = &b.buf[0][uint64_t(y) * b.bw];
			__m128i sum = _mm_mullo_epi32(ci_from_color64(line[x]), _mm_set1_epi32(ki[radius]));
			for(int d = 1; d <= radius; ++d)
			{
				const int32_t x0 = std::max(int32_t(x) - d, 0), x1 = std::min(int32_t(x) + d, last_x);
				const __m128i two = _mm_add_epi32(ci_from_color64(line[x0]), ci_from_color64(line[x1]));
				sum = _mm_add_epi32(sum, _mm_mullo_epi32(two, _mm_set1_epi32(ki[d + radius])));
			}
			b.buf[1][uint64_t(y) * b.bw + x] = ci_into_color64(_mm_srai_epi32(sum, color_shift));
		}
	for(int i = 0; i < MAX_BLOOM_KERNEL; ++i)
		ki[i] = int32_t(kernel[i] * bloom_scale * float(1 << color_shift));
	for(uint32_t y = 0; y < b.bh; ++y)
		for(uint32_t x = 0; x < b.bw; ++x)
		{
			const uint64_t* col = &b.buf[1][x];
			__m128i sum = _mm_mullo_epi32(ci_from_color64(col[uint64_t(y) * b.bw]), _mm_set1_epi32(ki[radius]));
			for(int d = 1; d <= radius; ++d)
			{
				const int32_t y0 = std::max(int32_t(y) - d, 0), y1 = std::min(int32_t(y) + d, last_y);
				const __m128i two =
					_mm_add_epi32(ci_from_color64(col[uint64_t(y0) * b.bw]), ci_from_color64(col[uint64_t(y1) * b.bw]));
				sum = _mm_add_epi32(sum, _mm_mullo_epi32(two, _mm_set1_epi32(ki[d + radius])));
			}
			b.buf[0][uint64_t(y) * b.bw + x] = ci_into_color64(_mm_srai_epi32(sum, color_shift));
		}
}

struct Tonemapper
{
	__m128 inv_scale, inv_255;
	Tonemapper(float exposure, const float modulate[3])
	{
		const float inv_255_s = 1.0f / 255.0f;
		const float is[3] = {1.0f / (exposure * modulate[0]), 1.0f / (exposure * modulate[1]), 1.0f / (exposure * modulate[2])};
		const float i255[3] = {inv_255_s, inv_255_s, inv_255_s};
		inv_scale = cv_from_f32x3(is);
		inv_255 = cv_from_f32x3(i255);
	}
	uint32_t map(__m128 c) const { return cv_into_color32(_mm_div_ps(c, _mm_fmadd_ps(c, inv_255, inv_scale))); }
};

// perform_tonemapping_with_bloom_nearest (postprocessor.rs:346-470)
void bloom_tonemap_nearest(const uint64_t* hdr, const Bloom& b, const Tonemapper& t, uint32_t* dst, uint32_t pitch)
{
	const uint32_t S = b.scale;
	const uint32_t columns_left = b.w - b.bw * S, lines_left = b.h - b.bh * S;
	auto put = [&](uint32_t x, uint32_t y, __m128 bloom) {
		dst[uint64_t(y) * pitch + x] = t.map(_mm_add_ps(bloom, cv_from_color64(hdr[uint64_t(y) * b.w + x])));
	};
	for(uint32_t sy = 0; sy < b.bh; ++sy)
	{
		for(uint32_t sx = 0; sx < b.bw; ++sx)
		{
			const __m128 bc = cv_from_color64(b.buf[0][uint64_t(sy) * b.bw + sx]);
			for(uint32_t dy = 0; dy < S; ++dy)
				for(uint32_t dx = 0; dx < S; ++dx)
					put(sx * S + dx, sy * S + dy, bc);
		}
		if(columns_left > 0)
		{
			const __m128 bc = cv_from_color64(b.buf[0][uint64_t(sy) * b.bw + b.bw - 1]);
			for(uint32_t dy = 0; dy < S; ++dy)
				for(uint32_t dx = 0; dx < columns_left; ++dx)
					put(b.bw * S + dx, sy * S + dy, bc);
		}
	}
	if(lines_left > 0)
	{
		const uint64_t* last = &b.buf[0][uint64_t(b.bh - 1) * b.bw];
		for(uint32_t sx = 0; sx < b.bw; ++sx)
		{
			const __m128 bc = cv_from_color64(last[sx]);
			for(uint32_t dy = 0; dy < lines_left; ++dy)
				for(uint32_t dx = 0; dx < S; ++dx)
					put(sx * S + dx, b.bh * S + dy, bc);
		}
		if(columns_left > 0)
		{
			const __m128 bc = cv_from_color64(last[b.bw - 1]);
			for(uint32_t dy = 0; dy < lines_left; ++dy)
				for(uint32_t dx = 0; dx < columns_left; ++dx)
					put(b.bw * S + dx, b.bh * S + dy, bc);
		}
	}
}

// perform_tonemapping_with_bloom_linear (postprocessor.rs:472-726)
void bloom_tonemap_linear(const uint64_t* hdr, const Bloom& b, const Tonemapper& t, uint32_t* dst, uint32_t pitch)
{
	const uint32_t S = b.scale;
	float mix[8];
	const float sf = float(S);
	for(uint32_t i = 0; i < S; ++i)
		mix[i] = float((S + 1) & 1) / (2.0f * sf) + float(i) / sf;
	auto bl = [&](uint32_t x, uint32_t y) { return cv_from_color64(b.buf[0][uint64_t(y) * b.bw + x]); };
	auto hdr_at = [&](uint32_t x, uint32_t y) { return cv_from_color64(hdr[uint64_t(y) * b.w + x]); };
	auto fms = [](__m128 a, float s, __m128 c) { return _mm_fmadd_ps(a, _mm_set1_ps(s), c); };
	auto out = [&](uint32_t x, uint32_t y, __m128 sum) { dst[uint64_t(y) * pitch + x] = t.map(sum); };
	const uint32_t right_x = (b.bw - 1) * S + S / 2, right_n = b.w - right_x;
	const uint32_t upper_y = (b.bh - 1) * S + S / 2, upper_n = b.h - upper_y;
	// lower border
	for(uint32_t sx = 0; sx + 1 < b.bw; ++sx)
	{
		const __m128 b00 = bl(sx, 0), b10 = bl(sx + 1, 0);
		for(uint32_t dy = 0; dy < S / 2; ++dy)
			for(uint32_t dx = 0; dx < S; ++dx)
			{
				const uint32_t x = sx * S + S / 2 + dx;
				out(x, dy, fms(b00, 1.0f - mix[dx], fms(b10, mix[dx], hdr_at(x, dy))));
			}
	}
	// main rows
	for(uint32_t sy = 0; sy + 1 < b.bh; ++sy)
	{
		const uint32_t y0 = sy * S + S / 2;
		for(uint32_t dy = 0; dy < S; ++dy)
		{
			const float my = mix[dy], omy = 1.0f - my;
			const uint32_t y = y0 + dy;
			const __m128 left = fms(bl(0, sy), omy, _mm_mul_ps(bl(0, sy + 1), _mm_set1_ps(my)));
			for(uint32_t x = 0; x < S / 2; ++x)
				out(x, y, _mm_add_ps(left, hdr_at(x, y)));
			for(uint32_t sx = 0; sx + 1 < b.bw; ++sx)
			{
				const __m128 l = fms(bl(sx, sy), omy, _mm_mul_ps(bl(sx, sy + 1), _mm_set1_ps(my)));
				const __m128 r = fms(bl(sx + 1, sy), omy, _mm_mul_ps(bl(sx + 1, sy + 1), _mm_set1_ps(my)));
				for(uint32_t dx = 0; dx < S; ++dx)
				{
					const uint32_t x = sx * S + S / 2 + dx;
					out(x, y, fms(l, 1.0f - mix[dx], fms(r, mix[dx], hdr_at(x, y))));
				}
			}
			const __m128 right = fms(bl(b.bw - 1, sy), omy, _mm_mul_ps(bl(b.bw - 1, sy + 1), _mm_set1_ps(my)));
			for(uint32_t dx = 0; dx < right_n; ++dx)
				out(right_x + dx, y, _mm_add_ps(right, hdr_at(right_x + dx, y)));
		}
	}
	// upper border
	for(uint32_t sx = 0; sx + 1 < b.bw; ++sx)
	{
		const __m128 b00 = bl(sx, b.bh - 1), b10 = bl(sx + 1, b.bh - 1);
		for(uint32_t dy = 0; dy < upper_n; ++dy)
			for(uint32_t dx = 0; dx < S; ++dx)
			{
				const uint32_t x = sx * S + S / 2 + dx, y = upper_y + dy;
				out(x, y, fms(b00, 1.0f - mix[dx], fms(b10, mix[dx], hdr_at(x, y))));
			}
	}
	// corners
	auto corner = [&](uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, __m128 v) {
		for(uint32_t y = y0; y < y1; ++y)
			for(uint32_t x = x0; x < x1; ++x)
				out(x, y, _mm_add_ps(v, hdr_at(x, y)));
	};
	corner(0, S / 2, 0, S / 2, bl(0, 0));
	corner(right_x, b.w, 0, S / 2, bl(b.bw - 1, 0));
	corner(0, S / 2, upper_y, b.h, bl(0, b.bh - 1));
	corner(right_x, b.w, upper_y, b.h, bl(b.bw - 1, b.bh - 1));
}

// ---------------------------------------------------------------------------------------------
// Triangle-model preparation (triangle_models_rendering.rs:7-200, 324-401, 702-731; light.rs:296-340),
// with cgmath 0.18's f32 matrix arithmetic as the reference calls it: Matrix4 * Vector4 is
// a * v.x + b * v.y + c * v.z + d * v.w over the columns a..d, Matrix * Matrix takes that per column
// of the right operand, dot is x + y + z of the products, all unfused.
// ---------------------------------------------------------------------------------------------
struct CgMat4
{
	float c[4][4]; // c[col][row]
};
struct CgMat3
{
	float c[3][3];
};

inline float cg_dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cg_cross(Vec3 a, Vec3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

CgMat4 cg_mat4(const float* m)
{
	CgMat4 r;
	for(int c = 0; c < 4; ++c)
		for(int k = 0; k < 4; ++k)
			r.c[c][k] = m[c * 4 + k];
	return r;
}

// Matrix4 * Vector4, then truncate (matrix.rs:125-128 with w = 1)
Vec3 cg_transform(const CgMat4& m, Vec3 v, float w)
{
	float out[3];
	for(int k = 0; k < 3; ++k)
		out[k] = m.c[0][k] * v.x + m.c[1][k] * v.y + m.c[2][k] * v.z + m.c[3][k] * w;
	return v3(out[0], out[1], out[2]);
}
Vec3 cg_mul3(const CgMat3& m, Vec3 v)
{
	float out[3];
	for(int k = 0; k < 3; ++k)
		out[k] = m.c[0][k] * v.x + m.c[1][k] * v.y + m.c[2][k] * v.z;
	return v3(out[0], out[1], out[2]);
}
CgMat4 cg_mul44(const CgMat4& a, const CgMat4& b)
{
	CgMat4 r;
	for(int col = 0; col < 4; ++col)
		for(int k = 0; k < 4; ++k)
			r.c[col][k] = a.c[0][k] * b.c[col][0] + a.c[1][k] * b.c[col][1] + a.c[2][k] * b.c[col][2] + a.c[3][k] * b.c[col][3];
	return r;
}
CgMat3 cg_mul33(const CgMat3& a, const CgMat3& b)
{
	CgMat3 r;
	for(int col = 0; col < 3; ++col)
		for(int k = 0; k < 3; ++k)
			r.c[col][k] = a.c[0][k] * b.c[col][0] + a.c[1][k] * b.c[col][1] + a.c[2][k] * b.c[col][2];
	return r;
}

// get_normals_matrix (:722-731): Matrix3::from_cols of the model matrix's first three columns
// (truncated), transposed, inverted — cgmath's Matrix3::invert: det = a0.x (a1.y a2.z - a2.y a1.z)
// - a1.x (a0.y a2.z - a2.y a0.z) + a2.x (a0.y a1.z - a1.y a0.z); the inverse is
// from_cols(a1 x a2 / det, a2 x a0 / det, a0 x a1 / det) transposed; None (-> identity) at det 0.
CgMat3 get_normals_matrix(const CgMat4& m)
{
	CgMat3 axis;
	for(int c = 0; c < 3; ++c)
		for(int k = 0; k < 3; ++k)
			axis.c[c][k] = m.c[c][k];
	CgMat3 a; // transpose
	for(int c = 0; c < 3; ++c)
		for(int k = 0; k < 3; ++k)
			a.c[c][k] = axis.c[k][c];
	const Vec3 a0 = v3(a.c[0][0], a.c[0][1], a.c[0][2]), a1 = v3(a.c[1][0], a.c[1][1], a.c[1][2]), a2 = v3(a.c[2][0], a.c[2][1], a.c[2][2]);
	const float det = a0.x * (a1.y * a2.z - a2.y * a1.z) - a1.x * (a0.y * a2.z - a2.y * a0.z) + a2.x * (a0.y * a1.z - a1.y * a0.z);
	CgMat3 r;
	if(det == 0.0f)
	{
		for(int c = 0; c < 3; ++c)
			for(int k = 0; k < 3; ++k)
				r.c[c][k] = c == k ? 1.0f : 0.0f;
		return r;
	}
	const Vec3 cols[3] = {cg_cross(a1, a2), cg_cross(a2, a0), cg_cross(a0, a1)};
	CgMat3 f;
	for(int c = 0; c < 3; ++c)
	{
		f.c[c][0] = cols[c].x / det;
		f.c[c][1] = cols[c].y / det;
		f.c[c][2] = cols[c].z / det;
	}
	for(int c = 0; c < 3; ++c)
		for(int k = 0; k < 3; ++k)
			r.c[c][k] = f.c[k][c];
	return r;
}

// get_light_cube_light (light.rs:296-340)
void get_light_cube_light(const float cube[6][3], Vec3 n, float out[3])
{
	float t[3] = {0.0f, 0.0f, 0.0f};
	const float comps[3] = {n.x, n.y, n.z};
	for(int axis = 0; axis < 3; ++axis)
	{
		if(comps[axis] <= 0.0f)
			for(int i = 0; i < 3; ++i)
				t[i] += cube[axis * 2][i] * (-comps[axis]);
		else
			for(int i = 0; i < 3; ++i)
				t[i] += cube[axis * 2 + 1][i] * comps[axis];
	}
	out[0] = t[0];
	out[1] = t[1];
	out[2] = t[2];
}

// get_vertex_light (:702-720)
void get_vertex_light(const Numerics& nm, const sw_model_instance& in, Vec3 n, float out[3])
{
	const Vec3 nn = n * nm.inv_sqrt_fast(fmax_r(cg_dot(n, n), 0.00000001f));
	get_light_cube_light(in.light_cube, nn, out);
	for(int d = 0; d < 2; ++d)
	{
		const Vec3 v = v3(in.light_vector_scaled[d][0], in.light_vector_scaled[d][1], in.light_vector_scaled[d][2]);
		const float dot = fmax_r(cg_dot(nn, v), 0.0f);
		for(int i = 0; i < 3; ++i)
			out[i] += in.light_color[d][i] * dot;
	}
}

void put_vertex(sw_model_vertex& dst, Vec3 pos, const float tc[2], const sw_model_instance& in, const float light[3])
{
	dst.pos[0] = pos.x;
	dst.pos[1] = pos.y;
	dst.pos[2] = pos.z;
	dst.tc[0] = tc[0] * in.tc_scale[0] + in.tc_shift[0];
	dst.tc[1] = tc[1] * in.tc_scale[1] + in.tc_shift[1];
	for(int i = 0; i < 3; ++i)
		dst.light[i] = light[i];
}

// animate_and_transform_triangle_mesh_vertices (:7-200) for one instance; false on an index out
// of range (clamped, as the device kernel does).
bool animate_and_transform(const Numerics& nm, const sw_model_mesh& mesh, const sw_model_instance& in, sw_model_vertex* dst)
{
	bool ok = true;
	const CgMat4 model = cg_mat4(in.model_matrix), mv = cg_mat4(in.model_view_matrix);
	const CgMat3 normals = get_normals_matrix(model);
	const uint32_t f0 = in.frames[0], f1 = in.frames[1];
	const float lerp0 = fmin_r(fmax_r(in.lerp, 0.0f), 1.0f);
	const float lerp1 = 1.0f - lerp0;
	const bool perform_lerp = lerp0 > 0.01f && lerp0 < 0.99f;
	const uint32_t nv = mesh.num_vertices;
	float light[3];
	if(mesh.kind == SW_MODEL_NON_ANIMATED)
	{
		const float* v = static_cast<const float*>(mesh.vertices);
		for(uint32_t i = 0; i < nv; ++i, v += 8)
		{
			get_vertex_light(nm, in, cg_mul3(normals, v3(v[3], v[4], v[5])), light);
			put_vertex(dst[i], cg_transform(mv, v3(v[0], v[1], v[2]), 1.0f), v + 6, in, light);
		}
	}
	else if(mesh.kind == SW_MODEL_VERTEX_ANIMATED)
	{
		const float* tc = static_cast<const float*>(mesh.vertices);
		const float* d0 = mesh.frame_vertices + uint64_t(f0) * nv * 6;
		const float* d1 = mesh.frame_vertices + uint64_t(f1) * nv * 6;
		for(uint32_t i = 0; i < nv; ++i)
		{
			const float* a = d0 + uint64_t(i) * 6;
			const float* b = d1 + uint64_t(i) * 6;
			Vec3 pos, normal;
			if(perform_lerp)
			{
				pos = v3(a[0], a[1], a[2]) * lerp0 + v3(b[0], b[1], b[2]) * lerp1;
				normal = v3(a[3], a[4], a[5]) * lerp0 + v3(b[3], b[4], b[5]) * lerp1;
			}
			else
			{
				const float* s = lerp0 > lerp1 ? a : b;
				pos = v3(s[0], s[1], s[2]);
				normal = v3(s[3], s[4], s[5]);
			}
			get_vertex_light(nm, in, cg_mul3(normals, normal), light);
			put_vertex(dst[i], cg_transform(mv, pos, 1.0f), tc + uint64_t(i) * 2, in, light);
		}
	}
	else if(mesh.num_bones == 0)
	{
		// SkeletonAnimated without frame bones: the source vertices (:96-107)
		const float* v = static_cast<const float*>(mesh.vertices);
		for(uint32_t i = 0; i < nv; ++i, v += 10)
		{
			get_vertex_light(nm, in, cg_mul3(normals, v3(v[5], v[6], v[7])), light);
			put_vertex(dst[i], cg_transform(mv, v3(v[2], v[3], v[4]), 1.0f), v, in, light);
		}
	}
	else
	{
		const uint32_t nb = mesh.num_bones;
		std::vector<CgMat4> matrices(nb);
		std::vector<CgMat3> normals_matrices(nb);
		const float* fb0 = mesh.frame_bones + uint64_t(f0) * nb * 16;
		const float* fb1 = mesh.frame_bones + uint64_t(f1) * nb * 16;
		for(uint32_t b = 0; b < nb; ++b)
		{
			CgMat4 mat;
			const CgMat4 m0 = cg_mat4(fb0 + uint64_t(b) * 16), m1 = cg_mat4(fb1 + uint64_t(b) * 16);
			if(perform_lerp)
			{
				for(int c = 0; c < 4; ++c)
					for(int k = 0; k < 4; ++k)
						mat.c[c][k] = m0.c[c][k] * lerp0 + m1.c[c][k] * lerp1;
			}
			else
				mat = lerp0 > lerp1 ? m0 : m1;
			uint32_t parent = mesh.bone_parents[b];
			if(parent < nb && parent >= b)
			{
				ok = false;
				parent = nb;
			}
			matrices[b] = parent < nb ? cg_mul44(matrices[parent], mat) : mat;
		}
		const float weight_scale = 1.0f / 255.0f;
		CgMat3 nws;
		CgMat4 mvws;
		for(int c = 0; c < 3; ++c)
			for(int k = 0; k < 3; ++k)
				nws.c[c][k] = normals.c[c][k] * weight_scale;
		for(int c = 0; c < 4; ++c)
			for(int k = 0; k < 4; ++k)
				mvws.c[c][k] = mv.c[c][k] * weight_scale;
		for(uint32_t b = 0; b < nb; ++b)
		{
			normals_matrices[b] = cg_mul33(nws, get_normals_matrix(matrices[b]));
			matrices[b] = cg_mul44(mvws, matrices[b]);
		}
		const uint8_t* base = static_cast<const uint8_t*>(mesh.vertices);
		for(uint32_t i = 0; i < nv; ++i)
		{
			const float* v = reinterpret_cast<const float*>(base + uint64_t(i) * 40);
			const uint8_t* bones = base + uint64_t(i) * 40 + 32;
			auto bone = [&](int k) {
				uint32_t bi = bones[k * 2];
				if(bi >= nb)
				{
					ok = false;
					bi = nb - 1;
				}
				return bi;
			};
			const uint32_t i0 = bone(0);
			const float w0 = float(bones[1]);
			CgMat4 mat;
			CgMat3 nmat;
			for(int c = 0; c < 4; ++c)
				for(int k = 0; k < 4; ++k)
					mat.c[c][k] = matrices[i0].c[c][k] * w0;
			for(int c = 0; c < 3; ++c)
				for(int k = 0; k < 3; ++k)
					nmat.c[c][k] = normals_matrices[i0].c[c][k] * w0;
			for(int j = 1; j < 4; ++j)
				if(bones[j * 2 + 1] > 0)
				{
					const uint32_t ii = bone(j);
					const float wi = float(bones[j * 2 + 1]);
					for(int c = 0; c < 4; ++c)
						for(int k = 0; k < 4; ++k)
							mat.c[c][k] += matrices[ii].c[c][k] * wi;
					for(int c = 0; c < 3; ++c)
						for(int k = 0; k < 3; ++k)
							nmat.c[c][k] += normals_matrices[ii].c[c][k] * wi;
				}
			get_vertex_light(nm, in, cg_mul3(nmat, v3(v[5], v[6], v[7])), light);
			put_vertex(dst[i], cg_transform(mat, v3(v[2], v[3], v[4]), 1.0f), v, in, light);
		}
	}
	return ok;
}

// reject_triangle_model_back_faces (:324-341) and sort_model_triangles (:343-391): the stable
// variant of the reference's comparator (greater max z first; equal depths in mesh order).
uint32_t reject_and_sort(const sw_model_mesh& mesh, const sw_model_vertex* verts, uint16_t* out, bool& ok)
{
	std::vector<uint32_t> tris; // indices of visible triangles
	auto vidx = [&](uint32_t t, int k) {
		uint32_t i = mesh.triangles[uint64_t(t) * 3 + k];
		if(i >= mesh.num_vertices)
		{
			ok = false;
			i = mesh.num_vertices - 1;
		}
		return i;
	};
	auto pos = [&](uint32_t i) { return v3(verts[i].pos[0], verts[i].pos[1], verts[i].pos[2]); };
	for(uint32_t t = 0; t < mesh.num_triangles; ++t)
	{
		// get_triangle_plane (:393-401)
		const Vec3 v0 = pos(vidx(t, 0)), v1 = pos(vidx(t, 1)), v2 = pos(vidx(t, 2));
		const Vec3 vec = cg_cross(v1 - v0, v2 - v1);
		if(cg_dot(vec, v0) > 0.0f)
			tris.push_back(t);
	}
	auto max_z = [&](uint32_t t) { return fmax_r(fmax_r(pos(vidx(t, 0)).z, pos(vidx(t, 1)).z), pos(vidx(t, 2)).z); };
	std::stable_sort(tris.begin(), tris.end(), [&](uint32_t a, uint32_t b) { return max_z(a) > max_z(b); });
	for(size_t i = 0; i < tris.size(); ++i)
		for(int k = 0; k < 3; ++k)
			out[i * 3 + k] = uint16_t(vidx(tris[i], k));
	return uint32_t(tris.size());
}

// ---------------------------------------------------------------------------------------------
// Per-frame texture animation: make_turb_distortion (generative_texture_effect_turb.rs:89-138),
// apply_texture_layer (generative_texture_effect_layered.rs:162-381) and the layered effect's
// emissive layers (mix_surface_with_texture, surfaces.rs:903-971), level by level as the reference
// loops.
// ---------------------------------------------------------------------------------------------
template <class T>
void turb_level(const T* src, T* dst, int32_t w, int32_t h, const int32_t* shifts)
{
	const int32_t mask[2] = {w - 1, h - 1};
	for(int32_t y = 0; y < h; ++y)
	{
		T* dst_row = dst + int64_t(y) * w;
		const int32_t row_shift = shifts[y];
		for(int32_t x = 0; x < w; ++x)
		{
			const int32_t column_shift = shifts[x];
			const int32_t u = x + row_shift;
			const int32_t v = y + column_shift;
			dst_row[x] = src[(u & mask[0]) + (v & mask[1]) * w];
		}
	}
}

void texture_layer(sw_texel* dst, const sw_texture_level& dl, const sw_texel* src, const sw_texture_level& sl, const int32_t shift[2],
	const float light[3], uint32_t blend)
{
	// fast path (:174-183): a plain copy
	if(blend == SW_BLEND_NONE && dl.width == sl.width && dl.height == sl.height && shift[0] == 0 && shift[1] == 0 && light[0] == 1.0f &&
		light[1] == 1.0f && light[2] == 1.0f)
	{
		std::memcpy(dst, src, sizeof(sw_texel) * dl.width * dl.height);
		return;
	}
	bool modulate = false; // impl_1 (:231-256)
	for(int c = 0; c < 3; ++c)
		modulate |= light[c] < 0.98f || light[c] > 1.02f;
	const float light_scale = float(1 << 8);
	const float lv[3] = {light[0] * light_scale, light[1] * light_scale, light[2] * light_scale};
	const __m128i light_vec = ci_from_f32x3(lv);
	for(uint32_t dst_v = 0; dst_v < dl.height; ++dst_v)
	{
		sw_texel* dst_line = dst + uint64_t(dst_v) * dl.width;
		const int32_t src_v = rem_euclid(shift[1] + int32_t(dst_v), int32_t(sl.height));
		const sw_texel* src_line = src + uint64_t(uint32_t(src_v)) * sl.width;
		int32_t src_u = rem_euclid(shift[0], int32_t(sl.width));
		for(uint32_t x = 0; x < dl.width; ++x)
		{
			const sw_texel texel_value = src_line[src_u];
			sw_texel& d = dst_line[x];
			if(modulate)
			{
				const __m128i m = _mm_srai_epi32(_mm_mullo_epi32(ci_from_color32(texel_value.diffuse), light_vec), 8);
				switch(blend)
				{
				case SW_BLEND_NONE:
					d.diffuse = ci_into_color32(m);
					d.packed_normal_roughness = texel_value.packed_normal_roughness;
					break;
				case SW_BLEND_AVERAGE: d.diffuse = ci_into_color32(_mm_srai_epi32(_mm_add_epi32(m, ci_from_color32(d.diffuse)), 1)); break;
				case SW_BLEND_ADDITIVE: d.diffuse = ci_into_color32(_mm_add_epi32(m, ci_from_color32(d.diffuse))); break;
				case SW_BLEND_ALPHA_TEST:
					if(C32::test_alpha(texel_value.diffuse))
					{
						d.diffuse = ci_into_color32(m);
						d.packed_normal_roughness = texel_value.packed_normal_roughness;
					}
					break;
				case SW_BLEND_ALPHA_BLEND:
				{
					const int32_t alpha = C32::get_alpha(texel_value.diffuse);
					d.diffuse = ci_into_color32(_mm_srai_epi32(_mm_add_epi32(_mm_mullo_epi32(m, _mm_set1_epi32(alpha)),
																  _mm_mullo_epi32(ci_from_color32(d.diffuse), _mm_set1_epi32(255 - alpha))),
						8));
					break;
				}
				default: break;
				}
			}
			else
			{
				switch(blend)
				{
				case SW_BLEND_NONE: d = texel_value; break;
				case SW_BLEND_AVERAGE: d.diffuse = C32::average(d.diffuse, texel_value.diffuse); break;
				case SW_BLEND_ADDITIVE:
					d.diffuse = ci_into_color32(_mm_add_epi32(ci_from_color32(texel_value.diffuse), ci_from_color32(d.diffuse)));
					break;
				case SW_BLEND_ALPHA_TEST:
					if(C32::test_alpha(texel_value.diffuse))
						d = texel_value;
					break;
				case SW_BLEND_ALPHA_BLEND:
				{
					const int32_t alpha = C32::get_alpha(texel_value.diffuse);
					d.diffuse = ci_into_color32(
						_mm_srai_epi32(_mm_add_epi32(_mm_mullo_epi32(ci_from_color32(texel_value.diffuse), _mm_set1_epi32(alpha)),
										   _mm_mullo_epi32(ci_from_color32(d.diffuse), _mm_set1_epi32(255 - alpha))),
							8));
					break;
				}
				default: break;
				}
			}
			++src_u;
			if(src_u == int32_t(sl.width))
				src_u = 0;
		}
	}
}

void lite_layer(uint32_t* dst, const sw_texture_level& dl, const uint32_t* src, const sw_texture_level& sl, const int32_t shift[2],
	const float light[3], uint32_t blend)
{
	const float light_scale = float(1 << 8);
	const float lv[3] = {light[0] * light_scale, light[1] * light_scale, light[2] * light_scale};
	const __m128i light_vec = ci_from_f32x3(lv);
	for(uint32_t dst_v = 0; dst_v < dl.height; ++dst_v)
	{
		uint32_t* dst_line = dst + uint64_t(dst_v) * dl.width;
		const int32_t src_v = rem_euclid(shift[1] + int32_t(dst_v), int32_t(sl.height));
		const uint32_t* src_line = src + uint64_t(uint32_t(src_v)) * sl.width;
		int32_t src_u = rem_euclid(shift[0], int32_t(sl.width));
		for(uint32_t x = 0; x < dl.width; ++x)
		{
			const uint32_t texel_value = src_line[src_u];
			const __m128i m = _mm_srai_epi32(_mm_mullo_epi32(ci_from_color32(texel_value), light_vec), 8);
			uint32_t& d = dst_line[x];
			switch(blend)
			{
			case SW_BLEND_NONE: d = ci_into_color32(m); break;
			case SW_BLEND_AVERAGE: d = ci_into_color32(_mm_srai_epi32(_mm_add_epi32(m, ci_from_color32(d)), 1)); break;
			case SW_BLEND_ADDITIVE: d = ci_into_color32(_mm_add_epi32(m, ci_from_color32(d))); break;
			case SW_BLEND_ALPHA_TEST:
				if(C32::test_alpha(texel_value))
					d = ci_into_color32(m);
				break;
			case SW_BLEND_ALPHA_BLEND:
			{
				const int32_t alpha = C32::get_alpha(texel_value);
				d = ci_into_color32(_mm_srai_epi32(
					_mm_add_epi32(_mm_mullo_epi32(m, _mm_set1_epi32(alpha)), _mm_mullo_epi32(ci_from_color32(d), _mm_set1_epi32(255 - alpha))),
					8));
				break;
			}
			default: break;
			}
			++src_u;
			if(src_u == int32_t(sl.width))
				src_u = 0;
		}
	}
}

// make_skybox_side_texture{,64} and make_skybox_side_texture_mips (textures.rs:183-286)
template <class CT>
uint32_t skybox_side(const typename CT::T* image, uint32_t image_w, uint32_t image_h, float brightness, typename CT::T* out)
{
	const int32_t shift = 8;
	const float scale = float(1 << shift);
	const float bf = brightness * scale;
	const int32_t brightness_i = bf != bf ? 0 : (bf <= -2147483648.0f ? INT32_MIN : (bf >= 2147483648.0f ? INT32_MAX : int32_t(bf)));
	const int32_t bv[3] = {brightness_i, brightness_i, brightness_i};
	const __m128i brightness_vec = ci_from_i32x3(bv);
	const uint32_t size = std::min(image_w, image_h);
	for(uint32_t y = 0; y < size; ++y)
		for(uint32_t x = 0; x < size; ++x)
		{
			const __m128i c = CT::to_veci(image[x + image_w * y]);
			out[x + y * size] = CT::from_veci(_mm_srai_epi32(_mm_mullo_epi32(c, brightness_vec), shift));
		}
	typename CT::T* prev = out;
	uint32_t prev_size = size;
	for(int i = 1; i < 5; ++i)
	{
		const uint32_t mip_size = prev_size >> 1;
		if(mip_size == 0)
			break;
		typename CT::T* mip = prev + uint64_t(prev_size) * prev_size;
		for(uint32_t y = 0; y < mip_size; ++y)
		{
			const uint64_t src_offset0 = uint64_t(y * 2) * prev_size, src_offset1 = uint64_t(y * 2 + 1) * prev_size;
			for(uint32_t x = 0; x < mip_size; ++x)
			{
				const uint32_t src_x = x * 2;
				const __m128i p00 = CT::to_veci(prev[src_x + src_offset0]), p01 = CT::to_veci(prev[src_x + src_offset1]);
				const __m128i p10 = CT::to_veci(prev[src_x + 1 + src_offset0]), p11 = CT::to_veci(prev[src_x + 1 + src_offset1]);
				mip[uint64_t(y) * mip_size + x] =
					CT::from_veci(_mm_srai_epi32(_mm_add_epi32(_mm_add_epi32(p00, p01), _mm_add_epi32(p10, p11)), 2));
			}
		}
		prev = mip;
		prev_size = mip_size;
	}
	return size;
}

// ---------------------------------------------------------------------------------------------
// DepthRasterizer (rasterizer.rs:1295-1474), loop for loop: the edge walk of fill_polygon and the
// per-pixel inv_z += d_inv_z_dx / per-row line_inv_z += d_inv_z_dy stepping of fill_polygon_part.
// ---------------------------------------------------------------------------------------------
struct DepthRasterizer
{
	float* depth_buffer;
	int32_t width, height;

	void fill_polygon_part(bool depth_test, Fixed16 y_start, Fixed16 y_end, Fixed16 left_x_start, Fixed16 left_dx_dy,
		Fixed16 right_x_start, Fixed16 right_dx_dy, const float eq[3])
	{
		const int32_t y_start_int = std::max(fixed16_round_to_int(y_start), 0);
		const int32_t y_end_int = std::min(fixed16_round_to_int(y_end), height);
		const Fixed16 y_start_delta = int_to_fixed16(y_start_int) + FIXED16_HALF - y_start;
		Fixed16 x_left = left_x_start + fixed16_mul(y_start_delta, left_dx_dy) + FIXED16_HALF;
		Fixed16 x_right = right_x_start + fixed16_mul(y_start_delta, right_dx_dy) + FIXED16_HALF;
		float line_inv_z = float(y_start_int) * eq[1] + 0.5f * (eq[0] + eq[1]) + eq[2];
		for(int32_t y_int = y_start_int; y_int < y_end_int; ++y_int)
		{
			const int32_t x_start_int = std::max(fixed16_floor_to_int(x_left), 0);
			const int32_t x_end_int = std::min(fixed16_floor_to_int(x_right), width);
			if(x_start_int < x_end_int)
			{
				float* line = depth_buffer + int64_t(y_int) * width;
				float inv_z = line_inv_z + float(x_start_int) * eq[0];
				for(int32_t x = x_start_int; x < x_end_int; ++x)
				{
					line[x] = depth_test ? fmax_r(inv_z, line[x]) : inv_z;
					inv_z += eq[0];
				}
			}
			x_left += left_dx_dy;
			x_right += right_dx_dy;
			line_inv_z += eq[1];
		}
	}

	void fill_polygon(bool depth_test, const int32_t* v, uint32_t n, const float eq[3])
	{
		uint32_t lower = 0;
		Fixed16 min_y = v[1];
		for(uint32_t i = 0; i < n; ++i)
			if(v[i * 2 + 1] < min_y)
			{
				min_y = v[i * 2 + 1];
				lower = i;
			}
		uint32_t left = lower, right = lower;
		Fixed16 cur_y = min_y;
		for(;;)
		{
			uint32_t next_left = left + n - 1;
			if(next_left >= n)
				next_left -= n;
			uint32_t next_right = right + 1;
			if(next_right >= n)
				next_right -= n;
			const int32_t *lv = v + left * 2, *nlv = v + next_left * 2, *rv = v + right * 2, *nrv = v + next_right * 2;
			const Fixed16 dy_left = nlv[1] - lv[1], dy_right = nrv[1] - rv[1];
			const Fixed16 next_y = std::min(nlv[1], nrv[1]);
			if(dy_left > FIXED16_HALF && dy_right > FIXED16_HALF)
			{
				const Fixed16 dx_dy_left = fixed16_div(nlv[0] - lv[0], dy_left);
				const Fixed16 dx_dy_right = fixed16_div(nrv[0] - rv[0], dy_right);
				fill_polygon_part(depth_test, cur_y, next_y, lv[0] + fixed16_mul(dx_dy_left, cur_y - lv[1]), dx_dy_left,
					rv[0] + fixed16_mul(dx_dy_right, cur_y - rv[1]), dx_dy_right, eq);
			}
			else if(dy_left > 0 && dy_right > 0)
			{
				const int32_t cur_y_int = fixed16_round_to_int(cur_y), next_y_int = fixed16_round_to_int(next_y);
				if(cur_y_int < next_y_int)
				{
					const Fixed16 thin_line_y = int_to_fixed16(cur_y_int) + FIXED16_HALF;
					const Fixed16 x_start_left = lv[0] + fixed16_mul_div(thin_line_y - lv[1], nlv[0] - lv[0], dy_left);
					const Fixed16 x_start_right = rv[0] + fixed16_mul_div(thin_line_y - rv[1], nrv[0] - rv[0], dy_right);
					fill_polygon_part(depth_test, cur_y, next_y, x_start_left, 0, x_start_right, 0, eq);
				}
			}
			if(next_left == next_right)
				break;
			if(v[next_right * 2 + 1] < v[next_left * 2 + 1])
				right = next_right;
			else
				left = next_left;
			cur_y = next_y;
		}
	}
};

} // namespace

// ---------------------------------------------------------------------------------------------
// C ABI of the oracle (loaded with ctypes by the tests and bench.py)
// ---------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------
// generative_texture_effect_water.rs: the field step and the wavy texture, following the
// reference's row cases (first row, inner rows, last row; first and last columns wrapped)
// ---------------------------------------------------------------------------------------------
// f(offset, x-1, x+1, y-1, y+1, x, y) over the field in the reference's order (:325-417, :548-682)
template <class F>
void water_rows(uint32_t w, uint32_t h, F&& f)
{
	auto row = [&](uint32_t y, uint32_t up_row, uint32_t down_row) {
		const uint32_t line = y * w, end = line + w - 1;
		f(line, end, line + 1, up_row, down_row, 0u, y);
		for(uint32_t o = line + 1; o < end; ++o)
			f(o, o - 1, o + 1, up_row + (o - line), down_row + (o - line), o - line, y);
		f(end, end - 1, line, up_row + w - 1, down_row + w - 1, w - 1, y);
	};
	row(0, (h - 1) * w, w);
	for(uint32_t y = 1; y < h - 1; ++y)
		row(y, (y - 1) * w, (y + 1) * w);
	row(h - 1, (h - 2) * w, 0);
}

// Color32::get_average_4 (color.rs:51-71)
uint32_t color32_average_4(const uint32_t c[4])
{
	uint32_t sum[4] = {0, 0, 0, 0};
	for(int k = 0; k < 4; ++k)
	{
		sum[0] += c[k] & 0x000000FFu;
		sum[1] += c[k] & 0x0000FF00u;
		sum[2] += c[k] & 0x00FF0000u;
		sum[3] += (c[k] & 0xFF000000u) >> 2;
	}
	return (sum[0] >> 2) | ((sum[1] >> 2) & 0x0000FF00u) | ((sum[2] >> 2) & 0x00FF0000u) | (sum[3] & 0xFF000000u);
}

const float MIN_VALID_ROUGHNESS = 1.0f / (127.0f - 1.0f); // textures.rs:418

// build_texture_mip (textures.rs:309-362)
void texture_mip(const sw_texel* prev, uint32_t pw, uint32_t ph, sw_texel* out)
{
	const uint32_t w = pw >> 1, h = ph >> 1;
	for(uint32_t y = 0; y < h; ++y)
		for(uint32_t x = 0; x < w; ++x)
		{
			const sw_texel p[4] = {prev[(y * 2) * pw + x * 2], prev[(y * 2 + 1) * pw + x * 2], prev[(y * 2) * pw + x * 2 + 1],
				prev[(y * 2 + 1) * pw + x * 2 + 1]};
			const uint32_t c[4] = {p[0].diffuse, p[1].diffuse, p[2].diffuse, p[3].diffuse};
			Vec3 n[4];
			float r[4];
			for(int k = 0; k < 4; ++k)
				unpack_normal_roughness(p[k].packed_normal_roughness, n[k], r[k]);
			const Vec3 sum = v3(n[0].x + n[1].x + n[2].x + n[3].x, n[0].y + n[1].y + n[2].y + n[3].y, n[0].z + n[1].z + n[2].z + n[3].z);
			const float len = std::fmax(std::sqrt(sum.x * sum.x + sum.y * sum.y + sum.z * sum.z), 0.000001f);
			const float half_cos = len * 0.25f;
			const float deviation = (1.0f - half_cos) / (1.0f - 0.5f);
			const float rs = r[0] + r[1] + r[2] + r[3];
			const float dr = std::fmin(std::fmax(std::fma(rs, 0.25f, deviation), MIN_VALID_ROUGHNESS), 1.0f);
			sw_texel d;
			d.diffuse = color32_average_4(c);
			d.packed_normal_roughness = pack_normal_roughness(v3(sum.x / len, sum.y / len, sum.z / len), dr);
			out[y * w + x] = d;
		}
}

extern "C" {

int oracle_render_batch(
	int mode, const sw_map_desc* map, const sw_batch* batch, void* arena, void* framebuffer, uint32_t num_threads,
	sw_stats* stats, uint64_t* out_of_bounds)
{
	if(!map || !batch || !arena || !framebuffer)
		return -1;
	if(batch->color_bits == 32)
		return render_batch<C32>(mode, map, batch, arena, framebuffer, num_threads, stats, out_of_bounds);
	if(batch->color_bits == 64)
		return render_batch<C64>(mode, map, batch, arena, framebuffer, num_threads, stats, out_of_bounds);
	return -1;
}

// Surfaces only (arena output), for the surface-builder parity tests.
int oracle_build_surfaces(int mode, const sw_map_desc* map, const sw_batch* b, void* arena_v, sw_stats* stats)
{
	Numerics nm{mode};
	MapData m{map->texels, map->lite_texels, map->lightmaps_simple, map->lightmaps_directional, map->levels, nullptr};
	Counters counters;
	for(uint32_t vi = 0; vi < b->num_views; ++vi)
	{
		const sw_view& view = b->views[vi];
		if(b->color_bits == 32)
			build_view_surfaces<C32>(
				nm, m, *b, view, static_cast<uint32_t*>(arena_v), view.first_surface, view.first_surface + view.num_surfaces,
				&counters);
		else
			build_view_surfaces<C64>(
				nm, m, *b, view, static_cast<uint64_t*>(arena_v), view.first_surface, view.first_surface + view.num_surfaces,
				&counters);
	}
	if(stats)
	{
		std::memset(stats, 0, sizeof(*stats));
		stats->surface_texels = counters.surface_texels;
		stats->lightmap_elements = counters.lightmap_elements;
		stats->emissive_texels = counters.emissive_texels;
	}
	return 0;
}

int oracle_tonemap(
	const uint64_t* hdr, uint32_t width, uint32_t height, float exposure, const float modulate[3], uint32_t* dst,
	uint32_t dst_pitch_pixels, float out_sum[3])
{
	__m128 s;
	tonemap_rows(hdr, width, 0, height, exposure, modulate, dst, dst_pitch_pixels, &s);
	alignas(16) float v[4];
	_mm_store_ps(v, s);
	// lanes [B, G, R, A] -> RGB
	out_sum[0] = v[2];
	out_sum[1] = v[1];
	out_sum[2] = v[0];
	return 0;
}

// Postprocessing with bloom (postprocessor.rs:68-204 with bloom_sigma > 0): returns the blur
// radius (or -1 on a frame smaller than one bloom texel); out_average_rgb: the downscaled image's
// average colour, which the reference passes to update_exposure.
int oracle_bloom(
	const uint64_t* hdr, uint32_t width, uint32_t height, float exposure, const float modulate[3], float sigma,
	uint32_t scale_log2, float bloom_scale, uint32_t linear, uint32_t* dst, uint32_t dst_pitch_pixels, float out_average_rgb[3],
	int32_t out_kernel[2 * MAX_BLOOM_KERNEL])
{
	Bloom b;
	b.w = width;
	b.h = height;
	b.scale = 1u << std::min<uint32_t>(std::max<uint32_t>(scale_log2, 1), 3);
	b.bw = width / b.scale;
	b.bh = height / b.scale;
	if(b.bw == 0 || b.bh == 0)
		return -1;
	b.buf[0].assign(uint64_t(b.bw) * b.bh, 0);
	b.buf[1].assign(uint64_t(b.bw) * b.bh, 0);
	bloom_downscale(hdr, b);
	const float sigma_c = sigma / float(b.scale);
	const int radius = std::min(int(std::max(std::ceil(3.0f * sigma_c - 0.5f), 0.0f)), MAX_BLOOM_RADIUS);
	float kernel[MAX_BLOOM_KERNEL];
	gaussian_kernel(sigma_c, radius, kernel);
	if(out_kernel)
		for(int i = 0; i < MAX_BLOOM_KERNEL; ++i)
		{
			out_kernel[i] = int32_t(kernel[i] * float(1 << 14));
			out_kernel[MAX_BLOOM_KERNEL + i] = int32_t(kernel[i] * bloom_scale * float(1 << 14));
		}
	if(radius > 0)
		bloom_blur(b, kernel, radius, bloom_scale);
	const Tonemapper t(exposure, modulate);
	if(linear)
		bloom_tonemap_linear(hdr, b, t, dst, dst_pitch_pixels);
	else
		bloom_tonemap_nearest(hdr, b, t, dst, dst_pitch_pixels);
	alignas(16) float v[4];
	_mm_store_ps(v, b.average);
	out_average_rgb[0] = v[2];
	out_average_rgb[1] = v[1];
	out_average_rgb[2] = v[0];
	return radius;
}

// Triangle-model preparation for a batch of instances (all pointers host memory). Returns 0, or
// 1 when an index was out of range (clamped).
int oracle_prepare_models(int mode, const sw_model_mesh* meshes, uint32_t num_meshes, const sw_model_instance* instances,
	uint32_t num_instances, sw_model_vertex* out_vertices, uint16_t* out_triangles, uint32_t* out_counts)
{
	const Numerics nm{mode};
	bool ok = true;
	for(uint32_t i = 0; i < num_instances; ++i)
	{
		const sw_model_instance& in = instances[i];
		if(in.mesh >= num_meshes)
			return -1;
		const sw_model_mesh& mesh = meshes[in.mesh];
		sw_model_vertex* verts = out_vertices + in.first_vertex;
		ok &= animate_and_transform(nm, mesh, in, verts);
		out_counts[i] = reject_and_sort(mesh, verts, out_triangles + uint64_t(in.first_triangle) * 3, ok);
	}
	return ok ? 0 : 1;
}

// Shadow maps for the following oracle calls (arrays stay owned by the caller); num 0: none.
void oracle_set_light_shadows(const sw_light_shadow* shadows, uint32_t num, const float* texels)
{
	g_shadows = shadows;
	g_num_shadows = num;
	g_shadow_texels = texels;
}

// get_light_shadow_factor of one map for a light-relative vector (for the tests).
float oracle_shadow_factor(int mode, const sw_light_shadow* shadow, const float* texels, const float vec[3])
{
	const sw_light_shadow* s0 = g_shadows;
	const uint32_t n0 = g_num_shadows;
	const float* t0 = g_shadow_texels;
	g_shadows = shadow;
	g_num_shadows = 1;
	g_shadow_texels = texels;
	const float r = get_light_shadow_factor(Numerics{mode}, 0, v3(vec[0], vec[1], vec[2]));
	g_shadows = s0;
	g_num_shadows = n0;
	g_shadow_texels = t0;
	return r;
}

// Depth buffers: every target's draws in order (sw_render_depth's inputs, host memory).
int oracle_render_depth(const sw_depth_target* targets, uint32_t num_targets, const sw_depth_draw* draws, const int32_t* vdata,
	float* texels)
{
	for(uint32_t t = 0; t < num_targets; ++t)
	{
		DepthRasterizer r{texels + targets[t].offset, int32_t(targets[t].width), int32_t(targets[t].height)};
		for(uint32_t k = targets[t].first_draw; k < targets[t].first_draw + targets[t].num_draws; ++k)
			r.fill_polygon(draws[k].depth_test != 0, vdata + draws[k].first_vertex_int, draws[k].num_vertices, draws[k].depth_equation);
		if(targets[t].flags & SW_DEPTH_CIRCLE)
		{
			// make_shadow_map_circle (light.rs:130-154)
			const float inf = 1.0e24f;
			const int32_t size = int32_t(targets[t].width), radius = size / 2;
			float* data = texels + targets[t].offset;
			for(int32_t y = 0; y < size; ++y)
			{
				const int32_t dy = y - radius;
				const int32_t dx =
					radius - std::min(int32_t(std::ceil(std::sqrt(float(std::max(radius * radius - dy * dy, 0))))), radius);
				float* line = data + y * size;
				for(int32_t x = 0; x < dx; ++x)
					line[x] = inf;
				for(int32_t x = size - dx; x < size; ++x)
					line[x] = inf;
			}
		}
	}
	return 0;
}

// Texture operations on host copies of the map's texel buffers, in order (the levels as in
// sw_map_desc::levels).
int oracle_apply_texture_ops(const sw_texture_level* levels, uint32_t num_levels, sw_texel* texels, uint32_t* lite,
	const sw_texture_op* ops, uint32_t num_ops, const int32_t* turb_shifts)
{
	for(uint32_t i = 0; i < num_ops; ++i)
	{
		const sw_texture_op& op = ops[i];
		if(op.dst_level >= num_levels || op.src_level >= num_levels)
			return -1;
		const sw_texture_level& dl = levels[op.dst_level];
		const sw_texture_level& sl = levels[op.src_level];
		const bool lite_level = (dl.flags & SW_TEX_LITE_LEVEL) != 0;
		if(op.kind == SW_TEXOP_TURB)
		{
			if(lite_level)
				turb_level(lite + sl.offset, lite + dl.offset, int32_t(dl.width), int32_t(dl.height), turb_shifts + op.turb_shifts);
			else
				turb_level(texels + sl.offset, texels + dl.offset, int32_t(dl.width), int32_t(dl.height), turb_shifts + op.turb_shifts);
		}
		else if(op.kind == SW_TEXOP_LAYER)
			texture_layer(texels + dl.offset, dl, texels + sl.offset, sl, op.shift, op.light, op.blend);
		else
			lite_layer(lite + dl.offset, dl, lite + sl.offset, sl, op.shift, op.light, op.blend);
	}
	return 0;
}

// The portal textures' overlays (build_view_portals_textures, partial_renderer.rs:2205-2276): for
// each portal with a texture, mix_surface_with_texture (surfaces.rs:903-971) of the texture's level
// at the portal's mip over the portal's rect, in order.
int oracle_mix_portal_textures(const sw_map_desc* map, uint32_t color_bits, const sw_portal_mix* mixes, uint32_t num_mixes,
	void* texels, uint64_t num_texels)
{
	if(!map || (color_bits != 32 && color_bits != 64))
		return -1;
	MapData m{};
	m.lite_texels = map->lite_texels;
	m.levels = map->levels;
	for(uint32_t i = 0; i < num_mixes; ++i)
	{
		const sw_portal_mix& x = mixes[i];
		if(x.level >= map->num_levels || !(map->levels[x.level].flags & SW_TEX_LITE_LEVEL))
			return -1;
		if(x.dst_offset + uint64_t(x.size[0]) * x.size[1] > num_texels)
			return -1;
		sw_surface_desc s{};
		s.size[0] = x.size[0];
		s.size[1] = x.size[1];
		s.tc_start[0] = x.tc_start[0];
		s.tc_start[1] = x.tc_start[1];
		s.emissive_level = x.level;
		for(int c = 0; c < 3; ++c)
			s.emissive_light[c] = x.light[c];
		if(color_bits == 32)
			mix_surface_with_texture<C32>(m, s, x.blend, static_cast<uint32_t*>(texels) + x.dst_offset, nullptr);
		else
			mix_surface_with_texture<C64>(m, s, x.blend, static_cast<uint64_t*>(texels) + x.dst_offset, nullptr);
	}
	return 0;
}

// draw_mesh's triangle loop (partial_renderer.rs:3612-3716) over prepared models, with the
// reference's model polygon clipping (clipping.rs:203-385), in sw_expand_mesh_runs' output layout.
struct OracleMV
{
	float pos[3];
	float tc[2];
	float light[3];
};

// get_model_line_plane_intersection / get_model_line_line_intersection (clipping.rs:315-352)
static OracleMV oracle_mv_lerp(const OracleMV& v0, const OracleMV& v1, float dist0, float dist1, int dims)
{
	const float dist_sum = dist1 - dist0;
	const float k0 = dist0 / dist_sum;
	const float k1 = dist1 / dist_sum;
	OracleMV r{};
	for(int i = 0; i < dims; ++i)
		r.pos[i] = v0.pos[i] * k1 - v1.pos[i] * k0;
	for(int i = 0; i < 2; ++i)
		r.tc[i] = v0.tc[i] * k1 - v1.tc[i] * k0;
	for(int i = 0; i < 3; ++i)
		r.light[i] = v0.light[i] * k1 - v1.light[i] * k0;
	return r;
}

// clip_3d_model_polygon_by_plane (clipping.rs:259-313): keeps dot(plane.vec, pos) >= plane.dist
static size_t oracle_clip3(const std::vector<OracleMV>& in, const float plane[4], std::vector<OracleMV>& out)
{
	out.clear();
	auto dotv = [&](const OracleMV& v) { return (plane[0] * v.pos[0] + plane[1] * v.pos[1]) + plane[2] * v.pos[2]; };
	const OracleMV* prev = &in.back();
	float prev_dist = dotv(*prev);
	for(const OracleMV& v : in)
	{
		const float dist = dotv(v);
		if(dist > plane[3])
		{
			if(prev_dist < plane[3])
				out.push_back(oracle_mv_lerp(*prev, v, dotv(*prev) - plane[3], dotv(v) - plane[3], 3));
			out.push_back(v);
		}
		else if(dist == plane[3])
			out.push_back(v);
		else if(prev_dist > plane[3])
			out.push_back(oracle_mv_lerp(*prev, v, dotv(*prev) - plane[3], dotv(v) - plane[3], 3));
		prev = &v;
		prev_dist = dist;
	}
	return out.size();
}

// clip_2d_model_polygon (clipping.rs:203-257): keeps pos . (l.x, l.y) - l.z >= 0
static size_t oracle_clip2(const std::vector<OracleMV>& in, const float line[3], std::vector<OracleMV>& out)
{
	out.clear();
	auto dist_of = [&](const OracleMV& v) { return (v.pos[0] * line[0] + v.pos[1] * line[1]) - line[2]; };
	const OracleMV* prev = &in.back();
	float prev_dist = dist_of(*prev);
	for(const OracleMV& v : in)
	{
		const float dist = dist_of(v);
		if(dist > 0.0f)
		{
			if(prev_dist < 0.0f)
				out.push_back(oracle_mv_lerp(*prev, v, prev_dist, dist, 2));
			out.push_back(v);
		}
		else if(dist == 0.0f)
			out.push_back(v);
		else if(prev_dist > 0.0f)
			out.push_back(oracle_mv_lerp(*prev, v, prev_dist, dist, 2));
		prev = &v;
		prev_dist = dist;
	}
	return out.size();
}

// f32_to_fixed16 (fixed_math.rs): a saturating cast of x * 2^16, NaN -> 0
static int32_t oracle_f32_to_fixed16(float x)
{
	const float y = x * 65536.0f;
	if(std::isnan(y))
		return 0;
	if(y >= 2147483647.0f)
		return INT32_MAX;
	if(y <= -2147483648.0f)
		return INT32_MIN;
	return int32_t(y);
}

int oracle_expand_mesh_runs(const sw_mesh_run* runs, uint32_t num_runs, const sw_model_vertex* vertices, uint64_t num_vertices,
	const uint16_t* triangles, uint64_t num_triangles, const uint32_t* counts, uint32_t num_counts, int32_t* out_ints,
	uint64_t num_ints, uint8_t* out_fans)
{
	uint64_t int_base = 0, fan_base = 0;
	for(uint32_t ri = 0; ri < num_runs; ++ri)
	{
		const sw_mesh_run& R = runs[ri];
		if(R.count_index >= num_counts || R.num_planes3d > SW_MESH_MAX_PLANES3D || R.num_planes2d > 4 ||
			uint64_t(R.first_vertex) + R.num_vertices > num_vertices || uint64_t(R.first_triangle) + R.max_triangles > num_triangles)
			return -1;
		const uint32_t fan_max = 1u + R.num_planes3d + R.num_planes2d;
		if(int_base + uint64_t(R.max_triangles) * fan_max * 21u > num_ints)
			return -1;
		const uint32_t visible = std::min(counts[R.count_index], R.max_triangles);
		std::vector<OracleMV> a, b;
		for(uint32_t t = 0; t < R.max_triangles; ++t)
		{
			uint32_t fan = 0;
			if(t < visible)
			{
				a.clear();
				for(int k = 0; k < 3; ++k)
				{
					const uint32_t vi = triangles[(uint64_t(R.first_triangle) + t) * 3 + k];
					if(vi >= R.num_vertices)
						return -1;
					const sw_model_vertex& v = vertices[R.first_vertex + vi];
					a.push_back(OracleMV{{v.pos[0], v.pos[1], v.pos[2]}, {v.tc[0], v.tc[1]}, {v.light[0], v.light[1], v.light[2]}});
				}
				bool culled = false;
				for(uint32_t p = 0; p < R.num_planes3d && !culled; ++p)
				{
					culled = oracle_clip3(a, R.planes3d[p], b) < 3;
					std::swap(a, b);
				}
				if(!culled)
				{
					for(OracleMV& v : a)
					{
						// pos.truncate() / pos.z
						const float z = v.pos[2];
						v.pos[0] = v.pos[0] / z;
						v.pos[1] = v.pos[1] / z;
						v.pos[2] = 0.0f;
					}
					for(uint32_t p = 0; p < R.num_planes2d && !culled; ++p)
					{
						culled = oracle_clip2(a, R.planes2d[p], b) < 3;
						std::swap(a, b);
					}
				}
				if(!culled)
				{
					fan = uint32_t(a.size()) - 2u;
					if(fan > fan_max)
						return -1;
					for(uint32_t j = 0; j < fan; ++j)
					{
						const OracleMV* tri[3] = {&a[0], &a[j + 1], &a[j + 2]};
						int32_t* o = out_ints + int_base + (uint64_t(t) * fan_max + j) * 21u;
						for(int v = 0; v < 3; ++v)
						{
							o[v * 7 + 0] = oracle_f32_to_fixed16(tri[v]->pos[0]);
							o[v * 7 + 1] = oracle_f32_to_fixed16(tri[v]->pos[1]);
							o[v * 7 + 2] = oracle_f32_to_fixed16(tri[v]->tc[0]);
							o[v * 7 + 3] = oracle_f32_to_fixed16(tri[v]->tc[1]);
							for(int c = 0; c < 3; ++c)
								o[v * 7 + 4 + c] = oracle_f32_to_fixed16(tri[v]->light[c]);
						}
					}
				}
			}
			out_fans[fan_base + t] = uint8_t(fan);
		}
		int_base += uint64_t(R.max_triangles) * fan_max * 21u;
		fan_base += R.max_triangles;
	}
	return 0;
}

// One water step: the points (add_point_value, :68-76), then update_wave_field into field_old
// (:223-224; the caller swaps the buffers as the reference does).
int oracle_water_step(float* field, float* field_old, uint32_t log2w, uint32_t log2h, float attenuation, const sw_wave_point* pts,
	uint32_t num_points)
{
	const uint32_t w = 1u << log2w, h = 1u << log2h;
	for(uint32_t i = 0; i < num_points; ++i)
	{
		const uint32_t a = (pts[i].x & (w - 1)) + ((pts[i].y & (h - 1)) << log2w);
		field[a] += pts[i].value;
		field_old[a] -= pts[i].value;
	}
	float* dst = field_old;
	const float* src = field;
	water_rows(w, h, [&](uint32_t o, uint32_t xm, uint32_t xp, uint32_t ym, uint32_t yp, uint32_t, uint32_t) {
		const float sum = src[xm] + src[xp] + src[ym] + src[yp];
		const float val = dst[o];
		dst[o] = (sum * 0.5f - val) * attenuation;
	});
	return 0;
}

// generative_texture_effect_fire.rs, step_update_heat_map (:62-326) after its sources: update_heat_map
// in place (:357-468) in the reference's row order, then set_heat for the points (:80-86).
int oracle_fire_step(uint8_t* heat_map, uint32_t log2w, uint32_t log2h, float attenuation, uint32_t slow, const sw_heat_point* pts,
	uint32_t num_points)
{
	const uint32_t w = 1u << log2w, h = 1u << log2h;
	const uint32_t SHIFT = 20;
	const uint32_t attenuation_i = uint32_t(std::min(std::max(attenuation, 0.0f), 1.0f) * float(1u << (SHIFT - 2)));
	auto f = [&](uint32_t o, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
		const uint32_t sum = uint32_t(heat_map[a]) + uint32_t(heat_map[b]) + uint32_t(heat_map[c]) + uint32_t(heat_map[d]);
		heat_map[o] = uint8_t((sum * attenuation_i) >> SHIFT);
	};
	if(slow)
	{
		for(uint32_t y = 0; y < h - 1; ++y)
		{
			const uint32_t ls = y * w, le = ls + w - 1;
			f(ls, le, ls, ls + 1, ls + w);
			for(uint32_t x = ls + 1; x < le; ++x)
				f(x, x - 1, x, x + 1, x + w);
			f(le, le - 1, le, ls, le + w);
		}
	}
	else
	{
		for(uint32_t y = 0; y < h - 2; ++y)
		{
			const uint32_t ls = y * w, le = ls + w - 1;
			f(ls, le + w, ls + w, ls + 1 + w, ls + w * 2);
			for(uint32_t x = ls + 1; x < le; ++x)
				f(x, x - 1 + w, x + w, x + 1 + w, x + w * 2);
			f(le, le - 1 + w, le + w, ls + w, le + w * 2);
		}
		const uint32_t ls = (h - 2) * w, le = ls + w - 1;
		f(ls, le + w, ls + w, ls + 1 + w, ls + w);
		for(uint32_t x = ls + 1; x < le; ++x)
			f(x, x - 1 + w, x + w, x + 1 + w, x + w);
		f(le, le - 1 + w, le + w, ls + w, le + w);
	}
	for(uint32_t x = 0; x < w; ++x)
		heat_map[(h - 1) * w + x] = 0;
	for(uint32_t i = 0; i < num_points; ++i)
	{
		uint8_t& v = heat_map[(pts[i].x & (w - 1)) + ((pts[i].y & (h - 1)) << log2w)];
		v = std::max(v, uint8_t(pts[i].heat));
	}
	return 0;
}

// generate_emissive_texture_based_on_heat_map (:470-486) into the lite level dst_level, then
// build_texture_lite_mip (textures.rs:385-415) into the next four levels.
int oracle_fire_make_texture(const sw_texture_level* levels, uint32_t num_levels, uint32_t* lite, const uint8_t* heat_map, uint32_t log2w,
	uint32_t log2h, uint32_t dst_level, const uint32_t* palette)
{
	if(uint64_t(dst_level) + 4 >= num_levels)
		return -1;
	const uint32_t w = 1u << log2w, h = 1u << log2h;
	uint32_t* dst = lite + levels[dst_level].offset;
	for(uint32_t i = 0; i < w * h; ++i)
		dst[i] = palette[heat_map[i]];
	for(uint32_t mip = 1; mip <= 4; ++mip)
	{
		const uint32_t pw = w >> (mip - 1), mw = pw >> 1, mh = (h >> (mip - 1)) >> 1;
		const uint32_t* prev = lite + levels[dst_level + mip - 1].offset;
		uint32_t* out = lite + levels[dst_level + mip].offset;
		for(uint32_t y = 0; y < mh; ++y)
			for(uint32_t x = 0; x < mw; ++x)
			{
				const uint32_t c[4] = {prev[(y * 2) * pw + x * 2], prev[(y * 2 + 1) * pw + x * 2], prev[(y * 2) * pw + x * 2 + 1],
					prev[(y * 2 + 1) * pw + x * 2 + 1]};
				out[y * mw + x] = color32_average_4(c);
			}
	}
	return 0;
}

// make_wavy_texture (:420-682) into the dst level, then its four mips (:293-299).
int oracle_water_make_texture(const sw_texture_level* levels, uint32_t num_levels, sw_texel* texels, const float* field, uint32_t log2w,
	uint32_t log2h, const sw_water_texture_params* p)
{
	if(uint64_t(p->dst_level) + 4 >= num_levels || p->color_level >= num_levels)
		return -1;
	const uint32_t w = 1u << log2w, h = 1u << log2h;
	sw_texel* dst = texels + levels[p->dst_level].offset;
	const sw_texel* color = texels + levels[p->color_level].offset;
	const float roughness = std::fmin(std::fmax(p->roughness, MIN_VALID_ROUGHNESS), 1.0f);
	// the first update's colour fill (:437-455)
	if(p->color_mode == SW_WATER_COLOR_KEEP && p->init == SW_WATER_INIT_BASE_COLOR)
		for(uint32_t i = 0; i < w * h; ++i)
			dst[i].diffuse = p->base_color;
	if(p->color_mode == SW_WATER_COLOR_KEEP && p->init == SW_WATER_INIT_SOURCE)
		for(uint32_t i = 0; i < w * h; ++i)
			dst[i].diffuse = color[i].diffuse;
	const float tc_deform_scale = 16.0f;
	water_rows(w, h, [&](uint32_t o, uint32_t xm, uint32_t xp, uint32_t ym, uint32_t yp, uint32_t x, uint32_t y) {
		const float dx = field[xp] - field[xm];
		const float dy = field[yp] - field[ym];
		// cgmath normalize: v * (1 / sqrt(x * x + y * y + z * z))
		const float k = 1.0f / std::sqrt(dx * dx + dy * dy + 1.0f * 1.0f);
		const Vec3 n = v3(dx * k, dy * k, 1.0f * k);
		dst[o].packed_normal_roughness = pack_normal_roughness(n, roughness);
		if(p->color_mode == SW_WATER_COLOR_DEFORMED)
		{
			const int32_t du = int32_t(n.x * tc_deform_scale), dv = int32_t(n.y * tc_deform_scale);
			const uint32_t u = uint32_t(int32_t(x) + du) & (w - 1), v = uint32_t(int32_t(y) + dv) & (h - 1);
			dst[o].diffuse = color[u + v * w].diffuse;
		}
		else if(p->color_mode == SW_WATER_COLOR_DEFORMED_X)
		{
			const int32_t du = int32_t(n.x * tc_deform_scale);
			const uint32_t u = uint32_t(int32_t(x) + du) & (w - 1);
			dst[o].diffuse = color[u + y * w].diffuse;
		}
	});
	for(uint32_t mip = 1; mip <= 4; ++mip)
		texture_mip(texels + levels[p->dst_level + mip - 1].offset, w >> (mip - 1), h >> (mip - 1), texels + levels[p->dst_level + mip].offset);
	return 0;
}

// A skybox side and its mips back to back into out (sized by the caller); returns the side's size.
uint32_t oracle_make_skybox_side(uint32_t color_bits, const void* image, uint32_t image_w, uint32_t image_h, float brightness, void* out)
{
	if(color_bits == 64)
		return skybox_side<C64>(static_cast<const uint64_t*>(image), image_w, image_h, brightness, static_cast<uint64_t*>(out));
	return skybox_side<C32>(static_cast<const uint32_t*>(image), image_w, image_h, brightness, static_cast<uint32_t*>(out));
}

// Elementary operations, exposed for the closed-form / exhaustive tests.
uint32_t oracle_color32_from_f32x4(const float bgra[4])
{
	return cv_into_color32(_mm_set_ps(bgra[3], bgra[2], bgra[1], bgra[0]));
}
uint64_t oracle_color64_from_f32x4(const float bgra[4])
{
	return cv_into_color64(_mm_set_ps(bgra[3], bgra[2], bgra[1], bgra[0]));
}
uint32_t oracle_blend32(uint32_t blend, uint32_t dst, uint32_t src)
{
	write_into_framebuffer<C32>(blend, dst, src);
	return dst;
}
uint64_t oracle_blend64(uint32_t blend, uint64_t dst, uint64_t src)
{
	write_into_framebuffer<C64>(blend, dst, src);
	return dst;
}
int32_t oracle_pack_normal_roughness(const float n[3], float roughness)
{
	return pack_normal_roughness(v3(n[0], n[1], n[2]), roughness);
}
void oracle_unpack_normal_roughness(int32_t packed, float out[4])
{
	Vec3 n;
	float r;
	unpack_normal_roughness(packed, n, r);
	out[0] = n.x;
	out[1] = n.y;
	out[2] = n.z;
	out[3] = r;
}
int32_t oracle_fixed16_mul(int32_t a, int32_t b) { return fixed16_mul(a, b); }
int32_t oracle_fixed16_div(int32_t a, int32_t b) { return fixed16_div(a, b); }
int32_t oracle_fixed16_mul_div(int32_t a, int32_t b, int32_t c) { return fixed16_mul_div(a, b, c); }
float oracle_specular_intensity(int mode, float angle_cos, float inv_roughness)
{
	return get_specular_intensity(Numerics{mode}, angle_cos, inv_roughness);
}
float oracle_fresnel_factor_base(float c) { return get_fresnel_factor_base(c); }
float oracle_specular_k_dielectric(float f, float r) { return get_specular_k_dielectric(f, r); }
float oracle_specular_k_metal(float f, float r) { return get_specular_k_metal(f, r); }

// Mip chain of a Texture (textures.rs:288-362): src is mip 0 (w x h), out receives mips 1..4
// back to back. Returns the number of texels written.
uint64_t oracle_build_texture_mips(const sw_texel* src, uint32_t w, uint32_t h, sw_texel* out)
{
	std::vector<sw_texel> prev(src, src + uint64_t(w) * h);
	uint32_t pw = w, ph = h;
	uint64_t written = 0;
	for(int mip = 1; mip <= 4; ++mip)
	{
		const uint32_t mw = pw >> 1, mh = ph >> 1;
		std::vector<sw_texel> cur(uint64_t(mw) * mh);
		for(uint32_t y = 0; y < mh; ++y)
			for(uint32_t x = 0; x < mw; ++x)
			{
				const sw_texel p00 = prev[(y * 2) * pw + x * 2];
				const sw_texel p01 = prev[(y * 2 + 1) * pw + x * 2];
				const sw_texel p10 = prev[(y * 2) * pw + x * 2 + 1];
				const sw_texel p11 = prev[(y * 2 + 1) * pw + x * 2 + 1];
				// color.rs:51-71
				uint32_t sum[4] = {0, 0, 0, 0};
				for(uint32_t c : {p00.diffuse, p01.diffuse, p10.diffuse, p11.diffuse})
				{
					sum[0] += c & 0x000000FFu;
					sum[1] += c & 0x0000FF00u;
					sum[2] += c & 0x00FF0000u;
					sum[3] += (c & 0xFF000000u) >> 2;
				}
				sw_texel d;
				d.diffuse = (sum[0] >> 2) | ((sum[1] >> 2) & 0x0000FF00u) | ((sum[2] >> 2) & 0x00FF0000u) | (sum[3] & 0xFF000000u);
				Vec3 n[4];
				float r[4];
				unpack_normal_roughness(p00.packed_normal_roughness, n[0], r[0]);
				unpack_normal_roughness(p01.packed_normal_roughness, n[1], r[1]);
				unpack_normal_roughness(p10.packed_normal_roughness, n[2], r[2]);
				unpack_normal_roughness(p11.packed_normal_roughness, n[3], r[3]);
				const Vec3 ns = ((n[0] + n[1]) + n[2]) + n[3];
				const float len = fmax_r(std::sqrt((ns.x * ns.x + ns.y * ns.y) + ns.z * ns.z), 0.000001f);
				const Vec3 dn = v3(ns.x / len, ns.y / len, ns.z / len);
				const float half_cos = len * 0.25f;
				const float deviation = (1.0f - half_cos) / (1.0f - 0.5f);
				const float rs = ((r[0] + r[1]) + r[2]) + r[3];
				const float min_valid = 1.0f / (127.0f - 1.0f);
				const float dr = fmin_r(fmax_r(f32_mul_add(rs, 0.25f, deviation), min_valid), 1.0f);
				d.packed_normal_roughness = pack_normal_roughness(dn, dr);
				cur[uint64_t(y) * mw + x] = d;
			}
		std::memcpy(out + written, cur.data(), cur.size() * sizeof(sw_texel));
		written += cur.size();
		prev.swap(cur);
		pw = mw;
		ph = mh;
	}
	return written;
}

} // extern "C"
