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
