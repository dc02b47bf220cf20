// The host side of the per-frame texture animation: the parameters the map materials processor
// evaluates on the CPU before the texel work (sw_apply_texture_ops) runs on the device.
#include <cmath>
#include <cstdint>

#include "host_internal.h"

namespace
{
const float TAU = 6.28318530717958647692f;

// Rust `f32 as i32`: truncation, saturating, NaN -> 0
int32_t f32_as_i32(float x)
{
	if(x != x)
		return 0;
	if(x <= -2147483648.0f)
		return INT32_MIN;
	if(x >= 2147483648.0f)
		return INT32_MAX;
	return int32_t(x);
}
} // namespace

extern "C" {

// make_turb_distortion's shift table (generative_texture_effect_turb.rs:98-112) in a buffer the
// caller keeps between calls, as the reference keeps its temp buffer: resized to max(width, height)
// (entries past the previous length zero, the rest kept), then entries [0, width) recomputed.
uint32_t swh_turb_shifts(
	float amplitude, float wave_length, float frequency, float time_s, uint32_t width, uint32_t height, uint32_t mip, int32_t* buffer,
	uint32_t previous_len)
{
	const uint32_t len = width > height ? width : height;
	for(uint32_t i = previous_len; i < len; ++i)
		buffer[i] = 0;
	const float mip_scale = 1.0f / float(1u << mip);
	const float amplitude_corrected = mip_scale * amplitude;
	const float frequency_scaled = TAU / (wave_length * mip_scale);
	const float time_based_shift = time_s * frequency * TAU;
	for(uint32_t x = 0; x < width; ++x)
		buffer[x] = f32_as_i32(std::round(std::sin(std::fma(float(x), frequency_scaled, time_based_shift)) * amplitude_corrected));
	return len;
}

// SingleArgumentFunction::evaluate (common/material_function.rs:90-156). kind: 0 constant (p[0]),
// 1 linear (scale, offset), 2 sine, 3 triangle, 4 saw-tooth, 5 step wave (frequency, phase,
// amplitude, offset).
float swh_evaluate_function(uint32_t kind, const float p[4], float x)
{
	switch(kind)
	{
	case 0: return p[0];
	case 1: return x * p[0] + p[1];
	case 2: return std::sin(x * p[0] * TAU + p[1]) * p[2] + p[3];
	case 3:
	{
		const float arg = x * p[0] + p[1];
		const float f = arg - std::floor(arg);
		const float wave = f <= 0.25f ? 4.0f * f : (f <= 0.75f ? 2.0f - 4.0f * f : -4.0f + 4.0f * f);
		return wave * p[2] + p[3];
	}
	case 4:
	{
		const float arg = x * p[0] + p[1];
		const float f = arg - std::floor(arg);
		return (f * 2.0f - 1.0f) * p[2] + p[3];
	}
	case 5:
	{
		const float arg = x * p[0] + p[1];
		const float f = arg - std::floor(arg);
		return (f <= 0.5f ? 1.0f : -1.0f) * p[2] + p[3];
	}
	default: return 0.0f;
	}
}

// The texture-coordinate shift of a layer at a mip: (f(t) as i32) >> mip
// (generative_texture_effect_layered.rs:78-80).
int32_t swh_layer_shift(float value, uint32_t mip) { return f32_as_i32(value) >> mip; }

} // extern "C"
