// Device-side numerics and colour model for the CDNA4 kernels.
//
// Semantics follow the reference's x86 SSE4.1/FMA layer (square_wheel/fast_math.rs:6-368,
// :730-798), restated for gfx950 VALU:
//   - f32_mul_add / _mm_fmadd_ps      -> fmaf (v_fma_f32); every other op is unfused
//                                        (this file is compiled with -ffp-contract=off)
//   - inv_sqrt_fast / inv_fast        -> correctly rounded 1/sqrt(x) and 1/x, i.e. the reference's
//                                        own non-SSE definitions (fast_math.rs:376-384)
//   - cvtps2dq after min(x, 255)      -> fminf (NaN -> 255 like minps), rintf (round half even),
//                                        explicit clamp to the i32 range (x86 "integer indefinite")
//   - packus / paddus / pavg          -> per-channel integer code below
//   - cvttss2si (32/64)               -> explicit range check, truncation
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SW_DEV __device__ __forceinline__
#define SW_HD __host__ __device__ __forceinline__

#include "sw_divide.h"

namespace swd
{

// ---------------------------------------------------------------------------------------------
// Scalar numerics
// ---------------------------------------------------------------------------------------------
SW_DEV float inv_sqrt_fast(float x) { return 1.0f / __builtin_sqrtf(x); }
SW_DEV float inv_fast(float x) { return 1.0f / x; }
SW_DEV float fmax_r(float a, float b) { return __builtin_fmaxf(a, b); }
SW_DEV float fmin_r(float a, float b) { return __builtin_fminf(a, b); }
SW_DEV float mul_add(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

const float MIN_POSITIVE_VALUE = 1.0f / float(1 << 30);

// x86 cvttss2si (64-bit): truncation, INT64_MIN when NaN or out of range (rasterizer.rs:1552-1556).
SW_DEV int64_t f32_to_i64_trunc_x86(float x)
{
	if(!(x >= -9223372036854775808.0f && x < 9223372036854775808.0f))
		return INT64_MIN;
	return (int64_t)x;
}
// x86 cvttss2si (32-bit) (fixed_math.rs:36-39).
SW_DEV int32_t f32_to_i32_trunc_x86(float x)
{
	if(!(x >= -2147483648.0f && x < 2147483648.0f))
		return INT32_MIN;
	return (int32_t)x;
}
// x86 cvtps2dq lane: round half to even, INT32_MIN when out of range.
SW_DEV int32_t f32_to_i32_rne_x86(float x)
{
	const float r = __builtin_rintf(x);
	if(!(r >= -2147483648.0f && r < 2147483648.0f))
		return INT32_MIN;
	return (int32_t)r;
}

struct V3
{
	float x, y, z;
};
SW_DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
SW_DEV V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
SW_DEV V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
SW_DEV V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
// fast_math.rs:780-798
SW_DEV float dot(V3 a, V3 b) { return mul_add(a.x, b.x, mul_add(a.y, b.y, a.z * b.z)); }
SW_DEV V3 scalar_mul_add(V3 a, float s, V3 b) { return v3(mul_add(a.x, s, b.x), mul_add(a.y, s, b.y), mul_add(a.z, s, b.z)); }
SW_DEV float len2(V3 v) { return dot(v, v); }

// ---------------------------------------------------------------------------------------------
// ColorVec: 4 float lanes in the reference's order [B, G, R, A] (fast_math.rs:98-248)
// ---------------------------------------------------------------------------------------------
struct CV
{
	float b, g, r, a;
};
SW_DEV CV cv(float b, float g, float r, float a) { return CV{b, g, r, a}; }
SW_DEV CV cv_zero() { return CV{0.0f, 0.0f, 0.0f, 0.0f}; }
// from_color_f32x3 / _with_one: input [R, G, B]
SW_DEV CV cv_rgb(float r, float g, float b) { return CV{b, g, r, 0.0f}; }
SW_DEV CV cv_rgb1(float r, float g, float b) { return CV{b, g, r, 1.0f}; }
SW_DEV CV cv_mul(CV x, CV y) { return CV{x.b * y.b, x.g * y.g, x.r * y.r, x.a * y.a}; }
SW_DEV CV cv_scalar_mul(CV x, float s) { return CV{x.b * s, x.g * s, x.r * s, x.a * s}; }
SW_DEV CV cv_mul_scalar_add(CV x, float s, CV y)
{
	return CV{mul_add(x.b, s, y.b), mul_add(x.g, s, y.g), mul_add(x.r, s, y.r), mul_add(x.a, s, y.a)};
}
SW_DEV CV cv_from_color32(uint32_t c)
{
	return CV{float(c & 0xFFu), float((c >> 8) & 0xFFu), float((c >> 16) & 0xFFu), float(c >> 24)};
}
SW_DEV CV cv_from_color64(uint64_t c)
{
	return CV{float(uint32_t(c) & 0xFFFFu), float(uint32_t(c >> 16) & 0xFFFFu), float(uint32_t(c >> 32) & 0xFFFFu),
		float(uint32_t(c >> 48))};
}
// fast_math.rs:131-144
// After the min the value is at most 255 (65535) and never NaN (the min returns 255 for a NaN, as
// minps does), so cvtps2dq's out-of-range result 0x80000000 can only come from values below -2^31:
// clamping those to -2^31 first gives the same low bits from an in-range conversion, without
// f32_to_i32_rne_x86's range test (equal for every float but signalling NaNs, which arithmetic
// never produces; tests/test_lane_conversion.py).
SW_DEV uint32_t lane_to_u8(float x) { return uint32_t(int32_t(__builtin_rintf(fmax_r(fmin_r(x, 255.0f), -2147483648.0f)))) & 0xFFu; }
SW_DEV uint32_t lane_to_u16(float x)
{
	return uint32_t(int32_t(__builtin_rintf(fmax_r(fmin_r(x, 65535.0f), -2147483648.0f)))) & 0xFFFFu;
}

// ---------------------------------------------------------------------------------------------
// ColorVecI lanes (i32) [B, G, R, A] and the two colour types
// ---------------------------------------------------------------------------------------------
struct CI
{
	int32_t b, g, r, a;
};
// _mm_packus_epi32 then _mm_packus_epi16 (the 16-bit value is read back as signed): fast_math.rs:317-326
SW_DEV uint32_t ci_lane_to_u8(int32_t x) { return (x > 0 && x < 32768) ? uint32_t(x < 255 ? x : 255) : 0u; }
// _mm_packus_epi32: fast_math.rs:328-336
SW_DEV uint32_t ci_lane_to_u16(int32_t x) { return x <= 0 ? 0u : (x >= 65535 ? 65535u : uint32_t(x)); }

struct Color32T
{
	typedef uint32_t T;
	static SW_DEV CV to_cv(T c) { return cv_from_color32(c); }
	static SW_DEV T from_cv(CV v) { return lane_to_u8(v.b) | (lane_to_u8(v.g) << 8) | (lane_to_u8(v.r) << 16) | (lane_to_u8(v.a) << 24); }
	static SW_DEV CI to_ci(T c)
	{
		return CI{int32_t(c & 0xFFu), int32_t((c >> 8) & 0xFFu), int32_t((c >> 16) & 0xFFu), int32_t(c >> 24)};
	}
	static SW_DEV T from_ci(CI v)
	{
		return ci_lane_to_u8(v.b) | (ci_lane_to_u8(v.g) << 8) | (ci_lane_to_u8(v.r) << 16) | (ci_lane_to_u8(v.a) << 24);
	}
	// pavgb, fast_math.rs:44-52: per byte (a + b + 1) >> 1
	static SW_DEV T average(T a, T b) { return (a | b) - (((a ^ b) >> 1) & 0x7F7F7F7Fu); }
	// paddusb, fast_math.rs:24-32
	static SW_DEV T saturated_sum(T a, T b)
	{
		uint32_t r = 0;
#pragma unroll
		for(int i = 0; i < 4; ++i)
		{
			const uint32_t s = ((a >> (8 * i)) & 0xFFu) + ((b >> (8 * i)) & 0xFFu);
			r |= (s > 0xFFu ? 0xFFu : s) << (8 * i);
		}
		return r;
	}
	static SW_DEV int32_t get_alpha(T c) { return int32_t(c >> 24); }
	static SW_DEV bool test_alpha(T c) { return (c & 0xFF000000u) > 0x7F000000u; }
};

struct Color64T
{
	typedef uint64_t T;
	static SW_DEV CV to_cv(T c) { return cv_from_color64(c); }
	static SW_DEV T from_cv(CV v)
	{
		return uint64_t(lane_to_u16(v.b)) | (uint64_t(lane_to_u16(v.g)) << 16) | (uint64_t(lane_to_u16(v.r)) << 32) |
			(uint64_t(lane_to_u16(v.a)) << 48);
	}
	static SW_DEV CI to_ci(T c)
	{
		return CI{int32_t(uint32_t(c) & 0xFFFFu), int32_t(uint32_t(c >> 16) & 0xFFFFu), int32_t(uint32_t(c >> 32) & 0xFFFFu),
			int32_t(uint32_t(c >> 48))};
	}
	static SW_DEV T from_ci(CI v)
	{
		return uint64_t(ci_lane_to_u16(v.b)) | (uint64_t(ci_lane_to_u16(v.g)) << 16) | (uint64_t(ci_lane_to_u16(v.r)) << 32) |
			(uint64_t(ci_lane_to_u16(v.a)) << 48);
	}
	// pavgw, fast_math.rs:54-62
	static SW_DEV T average(T a, T b) { return (a | b) - (((a ^ b) >> 1) & 0x7FFF7FFF7FFF7FFFull); }
	// paddusw, fast_math.rs:34-42
	static SW_DEV T saturated_sum(T a, T b)
	{
		uint64_t r = 0;
#pragma unroll
		for(int i = 0; i < 4; ++i)
		{
			const uint32_t s = uint32_t((a >> (16 * i)) & 0xFFFFu) + uint32_t((b >> (16 * i)) & 0xFFFFu);
			r |= uint64_t(s > 0xFFFFu ? 0xFFFFu : s) << (16 * i);
		}
		return r;
	}
	static SW_DEV int32_t get_alpha(T c) { return int32_t(uint32_t(c >> 48)); }
	static SW_DEV bool test_alpha(T c) { return (c & 0xFFFF000000000000ull) > 0x007F000000000000ull; }
};

// i32 lane ops with wrap-around (mullo_epi32 / add_epi32 / srai_epi32)
SW_DEV int32_t wmul(int32_t a, int32_t b) { return int32_t(uint32_t(a) * uint32_t(b)); }
SW_DEV int32_t wadd(int32_t a, int32_t b) { return int32_t(uint32_t(a) + uint32_t(b)); }

// abstract_color.rs:30-40 / :90-100 with the destination and source already unpacked
template <class CT>
SW_DEV typename CT::T alpha_blend_ci(typename CT::T dst, CI src, int32_t alpha)
{
	const CI d = CT::to_ci(dst);
	const int32_t ia = 255 - alpha;
	return CT::from_ci(CI{wadd(wmul(d.b, ia), wmul(src.b, alpha)) >> 8, wadd(wmul(d.g, ia), wmul(src.g, alpha)) >> 8,
		wadd(wmul(d.r, ia), wmul(src.r, alpha)) >> 8, wadd(wmul(d.a, ia), wmul(src.a, alpha)) >> 8});
}

// rasterizer.rs:1511-1536
template <class CT>
SW_DEV typename CT::T write_into_framebuffer(uint32_t blend, typename CT::T dst, typename CT::T texel)
{
	switch(blend)
	{
	case 0: return texel;
	case 1: return CT::average(dst, texel);
	case 2: return CT::saturated_sum(dst, texel);
	case 3: return CT::test_alpha(texel) ? texel : dst;
	case 4: return alpha_blend_ci<CT>(dst, CT::to_ci(texel), CT::get_alpha(texel));
	default: return dst;
	}
}

// ---------------------------------------------------------------------------------------------
// Fixed16 (common/fixed_math.rs)
// ---------------------------------------------------------------------------------------------
const int32_t FIXED16_ONE = 1 << 16;
const int32_t FIXED16_HALF = 1 << 15;
SW_DEV int32_t int_to_fixed16(int32_t x) { return int32_t(uint32_t(x) << 16); }
SW_DEV int32_t fixed16_round_to_int(int32_t x) { return wadd(x, FIXED16_HALF) >> 16; }
SW_DEV int32_t fixed16_floor_to_int(int32_t x) { return x >> 16; }
SW_DEV float fixed16_to_f32(int32_t x) { return float(x) / float(FIXED16_ONE); }
SW_DEV int32_t f32_to_fixed16_unchecked(float x) { return f32_to_i32_trunc_x86(x * float(FIXED16_ONE)); }
SW_DEV int32_t fixed16_mul(int32_t x, int32_t y) { return int32_t((int64_t(x) * int64_t(y)) >> 16); }
using swdiv::fixed16_div;
SW_DEV int32_t fixed16_mul_div(int32_t x, int32_t y, int32_t z)
{
	if(z == 0)
		return 0;
	return int32_t(int64_t(x) * int64_t(y) / int64_t(z));
}
SW_DEV int32_t idiv32(int32_t a, int32_t b)
{
	if(b == 0)
		return 0;
	if(a == INT32_MIN && b == -1)
		return INT32_MIN;
	return a / b;
}

const int TC_SHIFT = 24;
const int INV_Z_SHIFT = 29;
const int INV_Z_PRE_SHIFT = 8;
const int Z_CALC_SHIFT = 31;
const int TC_FINAL_SHIFT = TC_SHIFT + Z_CALC_SHIFT - INV_Z_SHIFT + INV_Z_PRE_SHIFT;
const float TC_SCALE = float(1 << TC_SHIFT);
const float INV_Z_SCALE = float(1 << INV_Z_SHIFT);

SW_DEV int32_t rem_euclid(int32_t a, int32_t b)
{
	const int32_t r = a % b;
	return r < 0 ? r + b : r;
}

} // namespace swd
