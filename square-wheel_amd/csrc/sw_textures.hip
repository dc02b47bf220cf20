// Per-frame texture animation on the map's resident textures (the map materials processor's
// generative effects): turbulence distortion (generative_texture_effect_turb.rs:89-138), texture
// layers (generative_texture_effect_layered.rs:162-381) and emissive layers (the layered effect's
// use of mix_surface_with_texture, surfaces.rs:854-971).
//
// One launch applies a list of operations. Consecutive operations with the same destination level
// form a run; every thread takes destination texels of one run and applies the run's operations to
// them in order (each operation reads only its source and the same destination texel, so doing
// them texel by texel gives the reference's level-by-level result).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sw_internal.h"
#include "sw_device.h"

using namespace swd;

namespace
{

#define TEXOP_BLOCK 256u
#define TEXOP_PER_THREAD 4u

struct TexLevel
{
	uint64_t offset;
	uint32_t w, h;
};

SW_DEV uint32_t wrap(int32_t v, uint32_t n)
{
	// rem_euclid for a positive divisor
	const int32_t r = v % int32_t(n);
	return uint32_t(r < 0 ? r + int32_t(n) : r);
}

// ColorVecI::from_color32(src) * light_vec >> 8 (light_vec = cvtps2dq(light * 256), alpha lane 0)
SW_DEV CI modulate(uint32_t c, const int32_t* lv)
{
	const CI t = Color32T::to_ci(c);
	return CI{wmul(t.b, lv[0]) >> 8, wmul(t.g, lv[1]) >> 8, wmul(t.r, lv[2]) >> 8, wmul(t.a, 0) >> 8};
}
SW_DEV CI ci_add2(CI a, CI b) { return CI{wadd(a.b, b.b), wadd(a.g, b.g), wadd(a.r, b.r), wadd(a.a, b.a)}; }
SW_DEV CI ci_shr2(CI a, int k) { return CI{a.b >> k, a.g >> k, a.r >> k, a.a >> k}; }
SW_DEV CI ci_blend(CI src, CI dst, int32_t alpha)
{
	const int32_t ia = 255 - alpha;
	return CI{wadd(wmul(src.b, alpha), wmul(dst.b, ia)) >> 8, wadd(wmul(src.g, alpha), wmul(dst.g, ia)) >> 8,
		wadd(wmul(src.r, alpha), wmul(dst.r, ia)) >> 8, wadd(wmul(src.a, alpha), wmul(dst.a, ia)) >> 8};
}

struct TexOpRun
{
	uint32_t first_op, num_ops;
	uint32_t first_block;
	uint32_t lite; // destination (and sources) are TextureLite levels
	TexLevel dst;
};

struct DevTexOp
{
	uint32_t kind, blend;
	uint32_t modulate; // layers: a light component outside [0.98, 1.02] (impl_1, :231-242)
	uint32_t turb_shifts; // TURB: first shift of the level
	int32_t shift[2];
	int32_t light_vec[3]; // lanes B, G, R
	TexLevel src;
};

struct TexOpLaunch
{
	const TexOpRun* runs;
	uint32_t num_runs;
	const DevTexOp* ops;
	const int32_t* turb_shifts;
	sw_texel* texels;
	uint32_t* lite;
};

// apply_texture_layer_impl_2 (generative_texture_effect_layered.rs:255-381), one texel
SW_DEV sw_texel layer_texel(const DevTexOp& op, sw_texel dst, sw_texel src)
{
	if(op.modulate)
	{
		const CI m = modulate(src.diffuse, op.light_vec);
		switch(op.blend)
		{
		case SW_BLEND_NONE: return sw_texel{Color32T::from_ci(m), src.packed_normal_roughness};
		case SW_BLEND_AVERAGE:
			dst.diffuse = Color32T::from_ci(ci_shr2(ci_add2(m, Color32T::to_ci(dst.diffuse)), 1));
			return dst;
		case SW_BLEND_ADDITIVE: dst.diffuse = Color32T::from_ci(ci_add2(m, Color32T::to_ci(dst.diffuse))); return dst;
		case SW_BLEND_ALPHA_TEST: return Color32T::test_alpha(src.diffuse) ? sw_texel{Color32T::from_ci(m), src.packed_normal_roughness} : dst;
		case SW_BLEND_ALPHA_BLEND:
			dst.diffuse = Color32T::from_ci(ci_blend(m, Color32T::to_ci(dst.diffuse), Color32T::get_alpha(src.diffuse)));
			return dst;
		default: return dst;
		}
	}
	switch(op.blend)
	{
	case SW_BLEND_NONE: return src;
	case SW_BLEND_AVERAGE: dst.diffuse = Color32T::average(dst.diffuse, src.diffuse); return dst;
	case SW_BLEND_ADDITIVE:
		dst.diffuse = Color32T::from_ci(ci_add2(Color32T::to_ci(src.diffuse), Color32T::to_ci(dst.diffuse)));
		return dst;
	case SW_BLEND_ALPHA_TEST: return Color32T::test_alpha(src.diffuse) ? src : dst;
	case SW_BLEND_ALPHA_BLEND:
		dst.diffuse =
			Color32T::from_ci(ci_blend(Color32T::to_ci(src.diffuse), Color32T::to_ci(dst.diffuse), Color32T::get_alpha(src.diffuse)));
		return dst;
	default: return dst;
	}
}

// mix_surface_with_texture_impl (surfaces.rs:903-971) on a Color32 destination, one texel
SW_DEV uint32_t lite_layer_texel(const DevTexOp& op, uint32_t dst, uint32_t src)
{
	const CI m = modulate(src, op.light_vec);
	switch(op.blend)
	{
	case SW_BLEND_NONE: return Color32T::from_ci(m);
	case SW_BLEND_AVERAGE: return Color32T::from_ci(ci_shr2(ci_add2(m, Color32T::to_ci(dst)), 1));
	case SW_BLEND_ADDITIVE: return Color32T::from_ci(ci_add2(m, Color32T::to_ci(dst)));
	case SW_BLEND_ALPHA_TEST: return Color32T::test_alpha(src) ? Color32T::from_ci(m) : dst;
	case SW_BLEND_ALPHA_BLEND: return Color32T::from_ci(ci_blend(m, Color32T::to_ci(dst), Color32T::get_alpha(src)));
	default: return dst;
	}
}

// make_turb_distortion (generative_texture_effect_turb.rs:89-138): the source texel at
// (x + shift[y], y + shift[x]), wrapped by the power-of-two masks
SW_DEV uint64_t turb_index(const TexOpLaunch& L, const DevTexOp& op, uint32_t x, uint32_t y, uint32_t w, uint32_t h)
{
	const int32_t row_shift = L.turb_shifts[op.turb_shifts + y];
	const int32_t column_shift = L.turb_shifts[op.turb_shifts + x];
	const uint32_t u = uint32_t(int32_t(x) + row_shift) & (w - 1), v = uint32_t(int32_t(y) + column_shift) & (h - 1);
	return op.src.offset + uint64_t(v) * w + u;
}

__global__ __launch_bounds__(TEXOP_BLOCK) void k_texture_ops(TexOpLaunch L)
{
	// the block's run: the last run starting at or before this block
	uint32_t lo = 0, hi = L.num_runs;
	while(hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if(L.runs[mid].first_block <= blockIdx.x)
			lo = mid;
		else
			hi = mid;
	}
	const TexOpRun run = L.runs[lo];
	const uint32_t w = run.dst.w, h = run.dst.h;
	const uint64_t n = uint64_t(w) * h;
	const uint64_t base = uint64_t(blockIdx.x - run.first_block) * (TEXOP_BLOCK * TEXOP_PER_THREAD);
	for(uint32_t k = 0; k < TEXOP_PER_THREAD; ++k)
	{
		const uint64_t i = base + k * TEXOP_BLOCK + threadIdx.x;
		if(i >= n)
			break;
		const uint32_t y = uint32_t(i / w), x = uint32_t(i - uint64_t(y) * w);
		if(run.lite)
		{
			uint32_t* d = L.lite + run.dst.offset + i;
			uint32_t v = *d;
			for(uint32_t o = 0; o < run.num_ops; ++o)
			{
				const DevTexOp& op = L.ops[run.first_op + o];
				if(op.kind == SW_TEXOP_TURB)
					v = L.lite[turb_index(L, op, x, y, w, h)];
				else
				{
					const uint32_t su = wrap(op.shift[0] + int32_t(x), op.src.w), sv = wrap(op.shift[1] + int32_t(y), op.src.h);
					v = lite_layer_texel(op, v, L.lite[op.src.offset + uint64_t(sv) * op.src.w + su]);
				}
			}
			*d = v;
		}
		else
		{
			sw_texel* d = L.texels + run.dst.offset + i;
			sw_texel v = *d;
			for(uint32_t o = 0; o < run.num_ops; ++o)
			{
				const DevTexOp& op = L.ops[run.first_op + o];
				if(op.kind == SW_TEXOP_TURB)
					v = L.texels[turb_index(L, op, x, y, w, h)];
				else
				{
					const uint32_t su = wrap(op.shift[0] + int32_t(x), op.src.w), sv = wrap(op.shift[1] + int32_t(y), op.src.h);
					v = layer_texel(op, v, L.texels[op.src.offset + uint64_t(sv) * op.src.w + su]);
				}
			}
			*d = v;
		}
	}
}

// make_skybox_side_texture / make_skybox_side_texture64 (textures.rs:183-229): the top-left
// size x size pixels of the image, (c * trunc(brightness * 256)) >> 8 per lane with the alpha lane
// multiplied by 0, packed with unsigned saturation.
template <class CT>
__global__ __launch_bounds__(256) void k_skybox_mip0(
	const typename CT::T* __restrict__ image, uint32_t image_w, uint32_t size, int32_t brightness, typename CT::T* __restrict__ dst)
{
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
	if(x >= size)
		return;
	const CI c = CT::to_ci(image[uint64_t(y) * image_w + x]);
	dst[uint64_t(y) * size + x] =
		CT::from_ci(CI{wmul(c.b, brightness) >> 8, wmul(c.g, brightness) >> 8, wmul(c.r, brightness) >> 8, wmul(c.a, 0) >> 8});
}

// make_skybox_side_texture_mips (textures.rs:231-286): (p00 + p01 + p10 + p11) >> 2 per lane
template <class CT>
__global__ __launch_bounds__(256) void k_skybox_mip(
	const typename CT::T* __restrict__ prev, uint32_t ps, uint32_t size, typename CT::T* __restrict__ dst)
{
	// prev: the previous level, ps x ps (ps = 2 size or 2 size + 1)
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
	if(x >= size)
		return;
	const CI p00 = CT::to_ci(prev[uint64_t(2 * y) * ps + 2 * x]), p01 = CT::to_ci(prev[uint64_t(2 * y + 1) * ps + 2 * x]);
	const CI p10 = CT::to_ci(prev[uint64_t(2 * y) * ps + 2 * x + 1]), p11 = CT::to_ci(prev[uint64_t(2 * y + 1) * ps + 2 * x + 1]);
	dst[uint64_t(y) * size + x] = CT::from_ci(ci_shr2(ci_add2(ci_add2(p00, p01), ci_add2(p10, p11)), 2));
}

} // namespace

namespace swk
{
// A skybox side and its mips (size, size/2, ... down to 1, at most five levels) back to back at dst.
void launch_skybox_side(hipStream_t st, uint32_t color_bits, const void* image, uint32_t image_w, uint32_t size, int32_t brightness,
	void* dst)
{
	if(color_bits == 64)
	{
		uint64_t* d = static_cast<uint64_t*>(dst);
		hipLaunchKernelGGL(k_skybox_mip0<Color64T>, dim3((size + 255) / 256, size), dim3(256), 0, st, static_cast<const uint64_t*>(image),
			image_w, size, brightness, d);
		for(uint32_t mip = 1, s = size; mip < 5 && (s >> 1); ++mip, d += uint64_t(s) * s, s >>= 1)
			hipLaunchKernelGGL(k_skybox_mip<Color64T>, dim3(((s >> 1) + 255) / 256, s >> 1), dim3(256), 0, st, d, s, s >> 1, d + uint64_t(s) * s);
	}
	else
	{
		uint32_t* d = static_cast<uint32_t*>(dst);
		hipLaunchKernelGGL(k_skybox_mip0<Color32T>, dim3((size + 255) / 256, size), dim3(256), 0, st, static_cast<const uint32_t*>(image),
			image_w, size, brightness, d);
		for(uint32_t mip = 1, s = size; mip < 5 && (s >> 1); ++mip, d += uint64_t(s) * s, s >>= 1)
			hipLaunchKernelGGL(k_skybox_mip<Color32T>, dim3(((s >> 1) + 255) / 256, s >> 1), dim3(256), 0, st, d, s, s >> 1, d + uint64_t(s) * s);
	}
}

// Validates the operations against the map's level table and packs the launch arguments (runs,
// then operations at `ops_at`) into `packed`.
int pack_texture_ops(const sw_texture_op* ops, uint32_t num_ops, const sw_texture_level* levels, uint32_t num_levels,
	uint32_t num_turb_shifts, std::vector<char>& packed, TexOpPlan& plan, std::string& error)
{
	std::vector<TexOpRun> runs;
	std::vector<DevTexOp> dops(num_ops);
	uint32_t blocks = 0;
	for(uint32_t i = 0; i < num_ops; ++i)
	{
		const sw_texture_op& op = ops[i];
		const std::string which = "texture op " + std::to_string(i) + ": ";
		if(op.kind > SW_TEXOP_LAYER_LITE)
			return (error = which + "unknown kind"), -1;
		if(op.dst_level >= num_levels || op.src_level >= num_levels)
			return (error = which + "level out of range"), -1;
		const sw_texture_level& dl = levels[op.dst_level];
		const sw_texture_level& sl = levels[op.src_level];
		const bool dlite = (dl.flags & SW_TEX_LITE_LEVEL) != 0, slite = (sl.flags & SW_TEX_LITE_LEVEL) != 0;
		if(dlite != slite || (op.kind == SW_TEXOP_LAYER && dlite) || (op.kind == SW_TEXOP_LAYER_LITE && !dlite))
			return (error = which + "level kinds do not fit the operation"), -1;
		if(op.dst_level == op.src_level)
			return (error = which + "source and destination are the same level"), -1;
		if(op.kind == SW_TEXOP_TURB)
		{
			if(sl.width != dl.width || sl.height != dl.height || (dl.width & (dl.width - 1)) || (dl.height & (dl.height - 1)))
				return (error = which + "turbulence needs equal power-of-two sizes"), -1;
			if(uint64_t(op.turb_shifts) + std::max(dl.width, dl.height) > num_turb_shifts)
				return (error = which + "turbulence shifts out of range"), -1;
		}
		else if(op.blend > SW_BLEND_ALPHA_BLEND)
			return (error = which + "unknown blending mode"), -1;
		DevTexOp& d = dops[i];
		d.kind = op.kind;
		d.blend = op.blend;
		d.modulate = 0;
		for(int c = 0; c < 3; ++c)
			d.modulate |= (op.light[c] < 0.98f || op.light[c] > 1.02f) ? 1u : 0u;
		d.turb_shifts = op.turb_shifts;
		d.shift[0] = op.shift[0];
		d.shift[1] = op.shift[1];
		// ColorVecI::from_color_f32x3(light * 256): cvtps2dq (round half even, "integer indefinite"
		// out of range), lanes B, G, R
		for(int c = 0; c < 3; ++c)
		{
			const float v = op.light[2 - c] * 256.0f;
			d.light_vec[c] = (v >= -2147483648.0f && v < 2147483648.0f) ? int32_t(std::nearbyint(v)) : INT32_MIN;
		}
		d.src = TexLevel{sl.offset, sl.width, sl.height};
		if(runs.empty() || ops[runs.back().first_op].dst_level != op.dst_level)
		{
			for(const TexOpRun& r : runs)
				if(ops[r.first_op].dst_level == op.dst_level)
					return (error = which + "a destination's operations must be consecutive"), -1;
			TexOpRun r{};
			r.first_op = i;
			r.num_ops = 0;
			r.first_block = blocks;
			r.lite = dlite ? 1u : 0u;
			r.dst = TexLevel{dl.offset, dl.width, dl.height};
			const uint64_t nb = (uint64_t(dl.width) * dl.height + TEXOP_BLOCK * TEXOP_PER_THREAD - 1) / (TEXOP_BLOCK * TEXOP_PER_THREAD);
			if(blocks + nb > 0x7FFFFFFFull)
				return (error = "texture ops: too many texels"), -1;
			blocks += uint32_t(nb);
			runs.push_back(r);
		}
		runs.back().num_ops++;
	}
	for(uint32_t i = 0; i < num_ops; ++i)
		for(const TexOpRun& r : runs)
			if(ops[i].src_level == ops[r.first_op].dst_level)
				return (error = "texture op " + std::to_string(i) + ": its source is the destination of another operation"), -1;
	const size_t runs_bytes = sizeof(TexOpRun) * runs.size();
	plan.ops_at = (runs_bytes + 255) & ~size_t(255);
	plan.bytes = plan.ops_at + sizeof(DevTexOp) * dops.size();
	plan.blocks = blocks;
	plan.num_runs = uint32_t(runs.size());
	packed.assign(plan.bytes, 0);
	std::memcpy(packed.data(), runs.data(), runs_bytes);
	std::memcpy(packed.data() + plan.ops_at, dops.data(), sizeof(DevTexOp) * dops.size());
	return 0;
}

void launch_texture_ops(hipStream_t st, const char* args_dev, const TexOpPlan& plan, const int32_t* turb_shifts_dev, sw_texel* texels,
	uint32_t* lite)
{
	if(plan.blocks == 0)
		return;
	TexOpLaunch L{reinterpret_cast<const TexOpRun*>(args_dev), plan.num_runs, reinterpret_cast<const DevTexOp*>(args_dev + plan.ops_at),
		turb_shifts_dev, texels, lite};
	hipLaunchKernelGGL(k_texture_ops, dim3(plan.blocks), dim3(TEXOP_BLOCK), 0, st, L);
}
} // namespace swk

namespace
{

// ---------------------------------------------------------------------------------------------
// Water (generative_texture_effect_water.rs)
// ---------------------------------------------------------------------------------------------
// add_point_value (:68-76) for the step's points, 64 at a time: the first point of each address in
// a batch adds the batch's values for that address in order (points at one address add up in the
// reference's order; distinct addresses go in parallel); batches follow each other.
__global__ __launch_bounds__(64) void k_water_points(float* field, float* field_old, const sw_wave_point* pts, uint32_t n,
	uint32_t log2w, uint32_t log2h)
{
	__shared__ uint32_t sa[64];
	__shared__ float sv[64];
	const uint32_t lane = threadIdx.x;
	const uint32_t mx = (1u << log2w) - 1u, my = (1u << log2h) - 1u;
	for(uint32_t base = 0; base < n; base += 64)
	{
		const uint32_t cnt = n - base < 64u ? n - base : 64u;
		uint32_t a = 0;
		if(lane < cnt)
		{
			const sw_wave_point p = pts[base + lane];
			a = (p.x & mx) + ((p.y & my) << log2w);
			sa[lane] = a;
			sv[lane] = p.value;
		}
		__syncthreads();
		if(lane < cnt)
		{
			bool first = true;
			for(uint32_t k = 0; k < lane; ++k)
				first = first && sa[k] != a;
			if(first)
			{
				float f = field[a], fo = field_old[a];
				for(uint32_t k = lane; k < cnt; ++k)
					if(sa[k] == a)
					{
						f += sv[k];
						fo -= sv[k];
					}
				field[a] = f;
				field_old[a] = fo;
			}
		}
		__syncthreads();
	}
}

// update_wave_field (:307-418): the field is a torus (rows and columns wrap); one lane per element
__global__ __launch_bounds__(256) void k_water_stencil(float* __restrict__ dst, const float* __restrict__ src, uint32_t log2w, uint32_t log2h, float attenuation)
{
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if(i >> (log2w + log2h))
		return;
	const uint32_t mx = (1u << log2w) - 1u, my = (1u << log2h) - 1u;
	const uint32_t x = i & mx, y = i >> log2w;
	const uint32_t row = y << log2w, up = ((y - 1u) & my) << log2w, down = ((y + 1u) & my) << log2w;
	const float sum = ((src[row + ((x - 1u) & mx)] + src[row + ((x + 1u) & mx)]) + src[up + x]) + src[down + x];
	dst[i] = (sum * 0.5f - dst[i]) * attenuation;
}

// PackedNormalRoughness::pack (textures.rs:33-39, fast_math.rs:64-78): x 127 per lane, cvtps2dq,
// signed saturation to bytes
SW_DEV int32_t pack_normal_roughness_dev(float x, float y, float z, float r)
{
	const float v[4] = {x * 127.0f, y * 127.0f, z * 127.0f, r * 127.0f};
	uint32_t out = 0;
	for(int k = 0; k < 4; ++k)
	{
		const int32_t q0 = f32_to_i32_rne_x86(v[k]);
		const int32_t q = q0 < -128 ? -128 : q0 > 127 ? 127 : q0;
		out |= (uint32_t(q) & 0xFFu) << (8 * k);
	}
	return int32_t(out);
}

// make_wavy_texture_impl (:497-682) per texel: the normal from the field's central differences,
// normalised as cgmath's normalize (v * (1 / |v|)), colours by mode
SW_DEV sw_texel water_texel(const float* field, uint32_t log2w, uint32_t log2h, const swk::WaterTexLaunch& w, const sw_texel* texels,
	uint32_t x, uint32_t y)
{
	const uint32_t mx = (1u << log2w) - 1u, my = (1u << log2h) - 1u;
	const uint32_t i = x + (y << log2w);
	const uint32_t row = y << log2w, up = ((y - 1u) & my) << log2w, down = ((y + 1u) & my) << log2w;
	const float dx = field[row + ((x + 1u) & mx)] - field[row + ((x - 1u) & mx)];
	const float dy = field[down + x] - field[up + x];
	const float m2 = (dx * dx + dy * dy) + 1.0f;
	const float s = 1.0f / __builtin_sqrtf(m2);
	const float nx = dx * s, ny = dy * s, nz = s;
	const sw_texel* color = texels + w.color_offset;
	uint32_t diffuse;
	if(w.color_mode == SW_WATER_COLOR_KEEP)
		diffuse = w.init == SW_WATER_INIT_BASE_COLOR ? w.base_color
			: w.init == SW_WATER_INIT_SOURCE      ? color[i].diffuse
												  : texels[w.dst_offset[0] + i].diffuse;
	else
	{
		const int32_t du = int32_t(nx * 16.0f);
		const uint32_t u = (x + uint32_t(du)) & mx;
		if(w.color_mode == SW_WATER_COLOR_DEFORMED)
		{
			const int32_t dv = int32_t(ny * 16.0f);
			const uint32_t v = (y + uint32_t(dv)) & my;
			diffuse = color[u + (v << log2w)].diffuse;
		}
		else
			diffuse = color[u + row].diffuse;
	}
	sw_texel t;
	t.diffuse = diffuse;
	t.packed_normal_roughness = pack_normal_roughness_dev(nx, ny, nz, w.roughness);
	return t;
}

// Color32::get_average_4 (color.rs:51-71)
SW_DEV uint32_t average_4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3)
{
	uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
	const uint32_t c[4] = {c0, c1, c2, c3};
	for(int k = 0; k < 4; ++k)
	{
		s0 += c[k] & 0x000000FFu;
		s1 += c[k] & 0x0000FF00u;
		s2 += c[k] & 0x00FF0000u;
		s3 += (c[k] & 0xFF000000u) >> 2;
	}
	return (s0 >> 2) | ((s1 >> 2) & 0x0000FF00u) | ((s2 >> 2) & 0x00FF0000u) | (s3 & 0xFF000000u);
}

SW_DEV void unpack_normal_roughness_dev(int32_t p, float& x, float& y, float& z, float& r)
{
	const float k = 1.0f / 127.0f;
	x = float(int8_t(p & 0xFF)) * k;
	y = float(int8_t((p >> 8) & 0xFF)) * k;
	z = float(int8_t((p >> 16) & 0xFF)) * k;
	r = float(int8_t((p >> 24) & 0xFF)) * k;
}

// build_texture_mip (textures.rs:309-362) for one texel
SW_DEV sw_texel texel_mip(sw_texel p00, sw_texel p01, sw_texel p10, sw_texel p11)
{
	float n[4][3], r[4];
	unpack_normal_roughness_dev(p00.packed_normal_roughness, n[0][0], n[0][1], n[0][2], r[0]);
	unpack_normal_roughness_dev(p01.packed_normal_roughness, n[1][0], n[1][1], n[1][2], r[1]);
	unpack_normal_roughness_dev(p10.packed_normal_roughness, n[2][0], n[2][1], n[2][2], r[2]);
	unpack_normal_roughness_dev(p11.packed_normal_roughness, n[3][0], n[3][1], n[3][2], r[3]);
	const float sx = ((n[0][0] + n[1][0]) + n[2][0]) + n[3][0];
	const float sy = ((n[0][1] + n[1][1]) + n[2][1]) + n[3][1];
	const float sz = ((n[0][2] + n[1][2]) + n[2][2]) + n[3][2];
	const float len = __builtin_fmaxf(__builtin_sqrtf((sx * sx + sy * sy) + sz * sz), 0.000001f);
	const float deviation = (1.0f - len * 0.25f) / 0.5f;
	const float rs = ((r[0] + r[1]) + r[2]) + r[3];
	const float dr = __builtin_fminf(__builtin_fmaxf(__builtin_fmaf(rs, 0.25f, deviation), 1.0f / 126.0f), 1.0f);
	sw_texel t;
	t.diffuse = average_4(p00.diffuse, p01.diffuse, p10.diffuse, p11.diffuse);
	t.packed_normal_roughness = pack_normal_roughness_dev(sx / len, sy / len, sz / len, dr);
	return t;
}

// The wavy texture of one 16 x 16 tile of mip 0 and the tile's texels of mips 1 .. 4, built in LDS
// (a mip texel depends only on the square of mip 0 under it): one launch for the texture and its mips.
__global__ __launch_bounds__(256) void k_water_texture_tiles(const float* field, uint32_t log2w, uint32_t log2h, swk::WaterTexLaunch w,
	sw_texel* texels)
{
	__shared__ sw_texel tile[16 * 16];
	const uint32_t tx = blockIdx.x & ((1u << (log2w - 4)) - 1u), ty = blockIdx.x >> (log2w - 4);
	const uint32_t lx = threadIdx.x & 15u, ly = threadIdx.x >> 4;
	const uint32_t x = tx * 16u + lx, y = ty * 16u + ly;
	const sw_texel t = water_texel(field, log2w, log2h, w, texels, x, y);
	tile[threadIdx.x] = t;
	texels[w.dst_offset[0] + x + (y << log2w)] = t; // each lane reads only its own old texel
	__syncthreads();
	for(uint32_t mip = 1, n = 8; mip <= 4; ++mip, n >>= 1)
	{
		sw_texel v{};
		const bool active = lx < n && ly < n;
		if(active)
			v = texel_mip(tile[(2u * ly) * 16u + 2u * lx], tile[(2u * ly + 1u) * 16u + 2u * lx], tile[(2u * ly) * 16u + 2u * lx + 1u],
				tile[(2u * ly + 1u) * 16u + 2u * lx + 1u]);
		__syncthreads();
		if(active)
		{
			tile[ly * 16u + lx] = v;
			texels[w.dst_offset[mip] + (tx * n + lx) + ((ty * n + ly) << (log2w - mip))] = v;
		}
		__syncthreads();
	}
}

// The palette texture (generate_emissive_texture_based_on_heat_map, :470-486) of one 16 x 16 tile and
// the tile's texels of lite mips 1 .. 4 (build_texture_lite_mip, textures.rs:385-415).
__global__ __launch_bounds__(256) void k_fire_texture_tiles(const uint8_t* heat, uint32_t log2w, const uint32_t* palette, uint32_t* lite,
	uint64_t o0, uint64_t o1, uint64_t o2, uint64_t o3, uint64_t o4)
{
	__shared__ uint32_t tile[16 * 16];
	const uint64_t off[5] = {o0, o1, o2, o3, o4};
	const uint32_t tx = blockIdx.x & ((1u << (log2w - 4)) - 1u), ty = blockIdx.x >> (log2w - 4);
	const uint32_t lx = threadIdx.x & 15u, ly = threadIdx.x >> 4;
	const uint32_t x = tx * 16u + lx, y = ty * 16u + ly;
	const uint32_t c = palette[heat[x + (y << log2w)]];
	tile[threadIdx.x] = c;
	lite[off[0] + x + (y << log2w)] = c;
	__syncthreads();
	for(uint32_t mip = 1, n = 8; mip <= 4; ++mip, n >>= 1)
	{
		uint32_t v = 0;
		const bool active = lx < n && ly < n;
		if(active)
			v = average_4(tile[(2u * ly) * 16u + 2u * lx], tile[(2u * ly + 1u) * 16u + 2u * lx], tile[(2u * ly) * 16u + 2u * lx + 1u],
				tile[(2u * ly + 1u) * 16u + 2u * lx + 1u]);
		__syncthreads();
		if(active)
		{
			tile[ly * 16u + lx] = v;
			lite[off[mip] + (tx * n + lx) + ((ty * n + ly) << (log2w - mip))] = v;
		}
		__syncthreads();
	}
}

} // namespace

namespace
{
// update_heat_map (:357-468), fast variant: row y < h - 2 from rows y + 1 (x - 1, x, x + 1, wrapped)
// and y + 2; row h - 2 takes row h - 1 twice; the last row is cleared. All reads see the old map
// (the reference walks rows downwards and reads only rows below), so one lane per element.
__global__ __launch_bounds__(256) void k_fire_update_fast(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint32_t log2w, uint32_t log2h, uint32_t att)
{
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if(i >> (log2w + log2h))
		return;
	const uint32_t w = 1u << log2w, h = 1u << log2h, mx = w - 1u;
	const uint32_t x = i & mx, y = i >> log2w;
	if(y == h - 1u)
	{
		dst[i] = 0;
		return;
	}
	const uint32_t r1 = (y + 1u) << log2w, r2 = (y + 2u < h ? y + 2u : y + 1u) << log2w;
	const uint32_t sum = uint32_t(src[r1 + ((x - 1u) & mx)]) + src[r1 + x] + src[r1 + ((x + 1u) & mx)] + src[r2 + x];
	dst[i] = uint8_t((sum * att) >> 20);
}

// the slow variant: each row is a chain (the left neighbour is the value just written; the last
// texel's right neighbour is the row's new first texel), from the row and the row below; one lane
// per row
__global__ __launch_bounds__(64) void k_fire_update_slow(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint32_t log2w, uint32_t log2h, uint32_t att)
{
	const uint32_t y = blockIdx.x * 64u + threadIdx.x;
	const uint32_t w = 1u << log2w, h = 1u << log2h;
	if(y >= h)
		return;
	const uint32_t row = y << log2w;
	if(y == h - 1u)
	{
		for(uint32_t x = 0; x < w; ++x)
			dst[row + x] = 0;
		return;
	}
	const uint32_t below = row + w;
	const uint32_t first = ((uint32_t(src[row + w - 1u]) + src[row] + src[row + 1u] + src[below]) * att) >> 20;
	dst[row] = uint8_t(first);
	uint32_t prev = first & 0xFFu;
	// the loads do not depend on the chain (src and dst are distinct buffers): unrolled, they overlap
#pragma unroll 8
	for(uint32_t x = 1; x < w - 1u; ++x)
	{
		const uint32_t v = ((prev + src[row + x] + src[row + x + 1u] + src[below + x]) * att) >> 20;
		dst[row + x] = uint8_t(v);
		prev = v & 0xFFu;
	}
	const uint32_t last = ((prev + src[row + w - 1u] + (first & 0xFFu) + src[below + w - 1u]) * att) >> 20;
	dst[row + w - 1u] = uint8_t(last);
}

// set_heat (:80-86) for the step's points, 64 at a time: the first point of each address takes the
// maximum of the batch's heats for it (max is order-free)
__global__ __launch_bounds__(64) void k_fire_points(uint8_t* heat, const sw_heat_point* pts, uint32_t n, uint32_t log2w, uint32_t log2h)
{
	__shared__ uint32_t sa[64], sh[64];
	const uint32_t lane = threadIdx.x;
	const uint32_t mx = (1u << log2w) - 1u, my = (1u << log2h) - 1u;
	for(uint32_t base = 0; base < n; base += 64)
	{
		const uint32_t cnt = n - base < 64u ? n - base : 64u;
		uint32_t a = 0;
		if(lane < cnt)
		{
			const sw_heat_point p = pts[base + lane];
			a = (p.x & mx) + ((p.y & my) << log2w);
			sa[lane] = a;
			sh[lane] = p.heat & 0xFFu;
		}
		__syncthreads();
		if(lane < cnt)
		{
			bool first = true;
			for(uint32_t k = 0; k < lane; ++k)
				first = first && sa[k] != a;
			if(first)
			{
				uint32_t v = heat[a];
				for(uint32_t k = lane; k < cnt; ++k)
					if(sa[k] == a && sh[k] > v)
						v = sh[k];
				heat[a] = uint8_t(v);
			}
		}
		__syncthreads();
	}
}

} // namespace

namespace swk
{
void launch_water_step(hipStream_t st, float* field, float* field_old, uint32_t log2w, uint32_t log2h, float attenuation,
	const sw_wave_point* points_dev, uint32_t num_points)
{
	if(num_points)
		hipLaunchKernelGGL(k_water_points, dim3(1), dim3(64), 0, st, field, field_old, points_dev, num_points, log2w, log2h);
	const uint32_t area = 1u << (log2w + log2h);
	hipLaunchKernelGGL(k_water_stencil, dim3((area + 255u) / 256u), dim3(256), 0, st, field_old, static_cast<const float*>(field), log2w,
		log2h, attenuation);
}

void launch_water_texture(hipStream_t st, const float* field, uint32_t log2w, uint32_t log2h, const WaterTexLaunch& w, sw_texel* texels)
{
	hipLaunchKernelGGL(k_water_texture_tiles, dim3(1u << (log2w + log2h - 8)), dim3(256), 0, st, field, log2w, log2h, w, texels);
}
void launch_fire_step(hipStream_t st, const uint8_t* heat, uint8_t* heat_old, uint32_t log2w, uint32_t log2h, uint32_t attenuation_i,
	bool slow, const sw_heat_point* points_dev, uint32_t num_points)
{
	const uint32_t area = 1u << (log2w + log2h), h = 1u << log2h;
	if(slow)
		hipLaunchKernelGGL(k_fire_update_slow, dim3((h + 63u) / 64u), dim3(64), 0, st, heat, heat_old, log2w, log2h, attenuation_i);
	else
		hipLaunchKernelGGL(k_fire_update_fast, dim3((area + 255u) / 256u), dim3(256), 0, st, heat, heat_old, log2w, log2h, attenuation_i);
	if(num_points)
		hipLaunchKernelGGL(k_fire_points, dim3(1), dim3(64), 0, st, heat_old, points_dev, num_points, log2w, log2h);
}

void launch_fire_texture(hipStream_t st, const uint8_t* heat, uint32_t log2w, uint32_t log2h, const uint64_t dst_offset[5],
	const uint32_t* palette_dev, uint32_t* lite)
{
	hipLaunchKernelGGL(k_fire_texture_tiles, dim3(1u << (log2w + log2h - 8)), dim3(256), 0, st, heat, log2w, palette_dev, lite, dst_offset[0],
		dst_offset[1], dst_offset[2], dst_offset[3], dst_offset[4]);
}

} // namespace swk
