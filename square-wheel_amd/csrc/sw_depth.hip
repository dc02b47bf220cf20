// The depth rasteriser of shadow maps (DepthRasterizer, rasterizer.rs:1295-1474): convex polygons
// with a 1 / z equation drawn into f32 depth buffers, overwriting (back-to-front order) or keeping
// the larger value (depth test).
//
// The reference steps 1 / z by additions along a span (inv_z += d_inv_z_dx) and along the rows of a
// polygon part (line_inv_z += d_inv_z_dy); float addition rounds at every step, so each value is the
// end of a chain of additions. The kernels keep those chains: a wave computes 64 consecutive values
// of a chain by a lane-serial scan (lane j adds to lane j - 1's value), chunk after chunk.
//
//   k_depth_setup   one wave per draw: the edge walk (lane 0) into parts, then per row of each part
//                   the span's x range and its first pixel's 1 / z (the rows' chain by the scan)
//   k_depth_raster  one wave per four depth-buffer rows, a 16-lane row of the wave per buffer row:
//                   the target's draws in order, each covering span filled 16 pixels at a time
//                   (the pixels' chain by a 16-lane scan, the four rows' chains side by side)
#include <cstring>
#include <string>
#include <vector>

#include "sw_internal.h"
#include "sw_device.h"

using namespace swd;

namespace
{

SW_DEV int32_t imax32(int32_t a, int32_t b) { return a > b ? a : b; }
SW_DEV int32_t imin32(int32_t a, int32_t b) { return a < b ? a : b; }

struct DepthRow
{
	uint16_t x0, x1; // [x0, x1): the span's pixels (empty when x0 >= x1)
	float inv_z; // 1 / z at x0
};

struct DepthDrawDev
{
	uint32_t first_vertex_int, num_vertices;
	float dx, dy, k; // DepthEquation
	uint32_t depth_test;
	uint32_t target;
	uint32_t pad;
	uint64_t rec_base; // the draw's row records: rec_base + y, y < the target's height
};

struct DepthTargetDev
{
	uint64_t offset; // of the depth buffer in the texels
	uint32_t w, h;
	uint32_t first_draw, num_draws;
	uint32_t first_row_block; // raster blocks: first_row_block + y / DEPTH_ROWS_PER_BLOCK
	uint32_t flags; // SW_DEPTH_*
};

// make_shadow_map_circle (light.rs:130-154): texels [0, dx) and [size - dx, size) of row y are
// outside the inscribed circle
SW_DEV uint32_t circle_margin(int32_t size, int32_t y)
{
	const int32_t radius = size / 2, dy = y - radius;
	const int32_t r2 = radius * radius - dy * dy;
	const int32_t half = int32_t(__builtin_ceilf(__builtin_sqrtf(float(r2 > 0 ? r2 : 0))));
	return uint32_t(radius - (half < radius ? half : radius));
}

struct DepthLaunch
{
	const DepthDrawDev* draws;
	const DepthTargetDev* targets;
	uint32_t num_targets;
	uint32_t pad;
	const int32_t* vdata;
	DepthRow* rows;
	int32_t* draw_rows; // per draw: [first, end) rows with records
	float* texels;
};

// value of lane j = value of lane j - 1 + d, lane 0 = start (a chain of j float additions). Each
// round moves every lane's value one lane up (the DPP wavefront shift wave_shr:1) and adds d; after
// round s lanes 0 .. s hold their final values. Called with all 64 lanes active.
SW_DEV float lane_shr1(float v)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
// within each 16-lane row of the wave: sublane j gets sublane j - 1's value (DPP row_shr:1), and
// the chain of four independent 16-value pieces (sublane 0 of each row takes its own start)
SW_DEV float row_shr1(float v)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xF, 0xF, false));
}
SW_DEV float serial_chain16(float start, float d)
{
	const bool first = (threadIdx.x & 15u) == 0;
	float v = start;
	for(uint32_t s = 1; s < 16; ++s)
	{
		const float p = row_shr1(v) + d;
		v = first ? start : p;
	}
	return v;
}
SW_DEV float serial_chain(float start, float d)
{
	const bool first = (threadIdx.x & 63u) == 0;
	float v = start;
	for(uint32_t s = 1; s < 64; ++s)
	{
		const float p = lane_shr1(v) + d;
		v = first ? start : p;
	}
	return v;
}

struct Part
{
	int32_t y0, y1; // rows [y0, y1), clamped to the buffer
	int32_t xl, xr; // x_left / x_right of row y0 (with + 0.5)
	int32_t dxl, dxr;
	float line_inv_z; // of row y0
};

#define MAX_DEPTH_PARTS SW_MAX_POLYGON_VERTICES

__global__ __launch_bounds__(64) void k_depth_setup(DepthLaunch L)
{
	__shared__ Part parts[MAX_DEPTH_PARTS];
	__shared__ uint32_t num_parts;
	const DepthDrawDev d = L.draws[blockIdx.x];
	const DepthTargetDev t = L.targets[d.target];
	const uint32_t lane = threadIdx.x;
	DepthRow* rows = L.rows + d.rec_base;
	for(uint32_t y = lane; y < t.h; y += 64)
		rows[y] = DepthRow{0, 0, 0.0f};
	if(lane == 0)
	{
		// fill_polygon (rasterizer.rs:1315-1412): from the lowest vertex up both chains
		const int32_t* v = L.vdata + d.first_vertex_int;
		const uint32_t n = d.num_vertices;
		uint32_t lower = 0;
		int32_t min_y = v[1];
		for(uint32_t i = 0; i < n; ++i)
			if(v[i * 2 + 1] < min_y)
			{
				min_y = v[i * 2 + 1];
				lower = i;
			}
		uint32_t left = lower, right = lower, np = 0;
		int32_t cur_y = min_y;
		const int32_t H = int32_t(t.h);
		for(uint32_t guard = 0; guard < 2 * n + 2; ++guard)
		{
			const uint32_t next_left = left + n - 1 >= n ? left + n - 1 - n : left + n - 1;
			const uint32_t next_right = right + 1 >= n ? right + 1 - n : right + 1;
			const int32_t lx = v[left * 2], ly = v[left * 2 + 1], nlx = v[next_left * 2], nly = v[next_left * 2 + 1];
			const int32_t rx = v[right * 2], ry = v[right * 2 + 1], nrx = v[next_right * 2], nry = v[next_right * 2 + 1];
			const int32_t dy_left = nly - ly, dy_right = nry - ry;
			const int32_t next_y = nly < nry ? nly : nry;
			bool fill = false;
			int32_t xs_l = 0, xs_r = 0, dl = 0, dr = 0;
			if(dy_left > FIXED16_HALF && dy_right > FIXED16_HALF)
			{
				dl = fixed16_div(nlx - lx, dy_left);
				dr = fixed16_div(nrx - rx, dy_right);
				xs_l = lx + fixed16_mul(dl, cur_y - ly);
				xs_r = rx + fixed16_mul(dr, cur_y - ry);
				fill = true;
			}
			else if(dy_left > 0 && dy_right > 0)
			{
				const int32_t cur_y_int = fixed16_round_to_int(cur_y), next_y_int = fixed16_round_to_int(next_y);
				if(cur_y_int < next_y_int)
				{
					const int32_t thin_line_y = int_to_fixed16(cur_y_int) + FIXED16_HALF;
					xs_l = lx + fixed16_mul_div(thin_line_y - ly, nlx - lx, dy_left);
					xs_r = rx + fixed16_mul_div(thin_line_y - ry, nrx - rx, dy_right);
					fill = true;
				}
			}
			if(fill && np < MAX_DEPTH_PARTS)
			{
				// fill_polygon_part (:1414-1465): the part's first row
				const int32_t y0 = imax32(fixed16_round_to_int(cur_y), 0);
				const int32_t y1 = imin32(fixed16_round_to_int(next_y), H);
				const int32_t y_start_delta = int_to_fixed16(y0) + FIXED16_HALF - cur_y;
				Part& p = parts[np++];
				p.y0 = y0;
				p.y1 = y1;
				p.xl = xs_l + fixed16_mul(y_start_delta, dl) + FIXED16_HALF;
				p.xr = xs_r + fixed16_mul(y_start_delta, dr) + FIXED16_HALF;
				p.dxl = dl;
				p.dxr = dr;
				p.line_inv_z = float(y0) * d.dy + 0.5f * (d.dx + d.dy) + d.k;
			}
			if(next_left == next_right)
				break;
			if(nry < nly)
				right = next_right;
			else
				left = next_left;
			cur_y = next_y;
		}
		num_parts = np;
	}
	__syncthreads();
	int32_t first = int32_t(t.h), end = 0;
	for(uint32_t pi = 0; pi < num_parts; ++pi)
	{
		const Part p = parts[pi];
		if(p.y0 >= p.y1)
			continue;
		first = imin32(first, p.y0);
		end = imax32(end, p.y1);
		float carry = p.line_inv_z;
		for(int32_t yb = p.y0; yb < p.y1; yb += 64)
		{
			const float line_inv_z = serial_chain(carry, d.dy);
			const int32_t y = yb + int32_t(lane);
			if(y < p.y1)
			{
				const uint32_t r = uint32_t(y - p.y0);
				const int32_t xl = int32_t(uint32_t(p.xl) + r * uint32_t(p.dxl)), xr = int32_t(uint32_t(p.xr) + r * uint32_t(p.dxr));
				const int32_t x0 = imax32(fixed16_floor_to_int(xl), 0), x1 = imin32(fixed16_floor_to_int(xr), int32_t(t.w));
				DepthRow row{0, 0, 0.0f};
				if(x0 < x1)
					row = DepthRow{uint16_t(x0), uint16_t(x1), line_inv_z + float(x0) * d.dx};
				rows[y] = row;
			}
			carry = __shfl(line_inv_z, 63) + d.dy;
		}
	}
	if(lane == 0)
	{
		L.draw_rows[blockIdx.x * 2] = first < end ? first : 0;
		L.draw_rows[blockIdx.x * 2 + 1] = first < end ? end : 0;
	}
}

#define DEPTH_ROWS_PER_BLOCK 4u // one 16-lane row of the wave per depth-buffer row
#define DEPTH_LDS_WIDTH 1024u // rows up to this width are drawn in LDS, written back once

__global__ __launch_bounds__(64) void k_depth_raster(DepthLaunch L)
{
	uint32_t lo = 0, hi = L.num_targets;
	while(hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if(L.targets[mid].first_row_block <= blockIdx.x)
			lo = mid;
		else
			hi = mid;
	}
	const DepthTargetDev t = L.targets[lo];
	const uint32_t lane = threadIdx.x, g = lane >> 4, sub = lane & 15u;
	const int32_t y = int32_t((blockIdx.x - t.first_row_block) * DEPTH_ROWS_PER_BLOCK + g);
	const bool row_ok = y < int32_t(t.h);
	__shared__ float srow[DEPTH_ROWS_PER_BLOCK * DEPTH_LDS_WIDTH];
	const bool in_lds = t.w <= DEPTH_LDS_WIDTH;
	float* grow = L.texels + t.offset + uint64_t(row_ok ? y : 0) * t.w;
	float* row = in_lds ? srow + g * DEPTH_LDS_WIDTH : grow;
	if(in_lds)
	{
		if(row_ok)
			for(uint32_t x = sub; x < t.w; x += 16)
				row[x] = grow[x];
		__syncthreads();
	}
	const uint32_t draws_end = t.first_draw + t.num_draws;
	for(uint32_t base = t.first_draw; base < draws_end; base += 16)
	{
		// lane (g, j) fetches draw base + j and its record of row y
		const uint32_t di = base + sub;
		uint32_t span = 0;
		float inv_z = 0.0f, dx = 0.0f;
		int32_t test = 0;
		if(di < draws_end)
		{
			const DepthDrawDev d = L.draws[di];
			dx = d.dx;
			test = int32_t(d.depth_test);
			const int32_t y_first = L.draw_rows[di * 2], y_end = L.draw_rows[di * 2 + 1];
			if(row_ok && y >= y_first && y < y_end)
			{
				const DepthRow r = L.rows[d.rec_base + uint32_t(y)];
				if(r.x0 < r.x1)
				{
					span = uint32_t(r.x0) | (uint32_t(r.x1) << 16);
					inv_z = r.inv_z;
				}
			}
		}
		const uint64_t any = __ballot(span != 0u);
		uint32_t draw_mask = uint32_t((any | (any >> 16) | (any >> 32) | (any >> 48)) & 0xFFFFu);
		while(draw_mask)
		{
			// the draws in order; each group draws its own row's span
			const int j = __builtin_ctz(draw_mask);
			draw_mask &= draw_mask - 1u;
			const float ddx = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, dx), j));
			const bool dtest = __builtin_amdgcn_readlane(test, j) != 0;
			const int src = int((g << 4) | uint32_t(j));
			const uint32_t sp = uint32_t(__shfl(int(span), src));
			float carry = __shfl(inv_z, src);
			const uint32_t x0 = sp & 0xFFFFu, len = (sp >> 16) - x0; // 0 where the row is not covered (sp = 0)
			const uint32_t len_all = max(max(uint32_t(__shfl(int(len), 0)), uint32_t(__shfl(int(len), 16))),
				max(uint32_t(__shfl(int(len), 32)), uint32_t(__shfl(int(len), 48))));
			for(uint32_t c = 0; c < len_all; c += 16)
			{
				const float v = serial_chain16(carry, ddx);
				if(c + sub < len)
				{
					const uint32_t x = x0 + c + sub;
					row[x] = dtest ? fmax_r(v, row[x]) : v;
				}
				carry = __shfl(v, int((g << 4) | 15u)) + ddx;
			}
		}
	}
	__syncthreads();
	if((t.flags & SW_DEPTH_CIRCLE) && row_ok)
	{
		const uint32_t dx = circle_margin(int32_t(t.w), y);
		for(uint32_t x = sub; x < t.w; x += 16)
			if(x < dx || x >= t.w - dx)
				row[x] = 1.0e24f;
	}
	if(in_lds && row_ok)
		for(uint32_t x = sub; x < t.w; x += 16)
			grow[x] = row[x];
}

} // namespace

namespace swk
{
int depth_plan(const sw_depth_target* targets, uint32_t num_targets, const sw_depth_draw* draws, uint32_t num_draws,
	uint64_t num_vertex_ints, uint64_t num_texels, std::vector<char>& packed, DepthPlan& plan, std::string& error)
{
	std::vector<DepthTargetDev> tg(num_targets);
	std::vector<DepthDrawDev> dd(num_draws);
	uint64_t rec = 0;
	uint32_t row_blocks = 0, next_draw = 0;
	for(uint32_t i = 0; i < num_targets; ++i)
	{
		const sw_depth_target& t = targets[i];
		const std::string which = "depth target " + std::to_string(i) + ": ";
		if(t.width == 0 || t.height == 0 || t.width > 65535 || t.height > 65535)
			return (error = which + "size out of range"), -1;
		if(t.offset > num_texels || uint64_t(t.width) * t.height > num_texels - t.offset)
			return (error = which + "buffer outside the depth texels"), -1;
		if(t.first_draw != next_draw || uint64_t(t.first_draw) + t.num_draws > num_draws)
			return (error = which + "draw range not following the previous target's"), -1;
		next_draw = t.first_draw + t.num_draws;
		if(t.flags & ~SW_DEPTH_CIRCLE)
			return (error = which + "unknown flags"), -1;
		if((t.flags & SW_DEPTH_CIRCLE) && t.width != t.height)
			return (error = which + "the circle needs a square target"), -1;
		if(uint64_t(row_blocks) + t.height > 0x7FFFFFFFu)
			return (error = "depth targets: too many rows"), -1;
		tg[i] = DepthTargetDev{t.offset, t.width, t.height, t.first_draw, t.num_draws, row_blocks, t.flags};
		row_blocks += (t.height + DEPTH_ROWS_PER_BLOCK - 1) / DEPTH_ROWS_PER_BLOCK;
		for(uint32_t k = t.first_draw; k < t.first_draw + t.num_draws; ++k)
		{
			const sw_depth_draw& s = draws[k];
			if(s.num_vertices < 3 || s.num_vertices > SW_MAX_POLYGON_VERTICES ||
				uint64_t(s.first_vertex_int) + 2ull * s.num_vertices > num_vertex_ints)
				return (error = "depth draw " + std::to_string(k) + ": vertices out of range"), -1;
			dd[k] = DepthDrawDev{s.first_vertex_int, s.num_vertices, s.depth_equation[0], s.depth_equation[1], s.depth_equation[2],
				s.depth_test ? 1u : 0u, i, 0, rec};
			rec += t.height;
		}
	}
	if(next_draw != num_draws)
		return (error = "depth draws not covered by the targets"), -1;
	plan.targets_at = (sizeof(DepthDrawDev) * dd.size() + 255) & ~size_t(255);
	plan.bytes = plan.targets_at + sizeof(DepthTargetDev) * tg.size();
	plan.row_records = rec;
	plan.row_blocks = row_blocks;
	plan.num_draws = num_draws;
	plan.num_targets = num_targets;
	packed.assign(plan.bytes, 0);
	std::memcpy(packed.data(), dd.data(), sizeof(DepthDrawDev) * dd.size());
	std::memcpy(packed.data() + plan.targets_at, tg.data(), sizeof(DepthTargetDev) * tg.size());
	return 0;
}

size_t depth_row_record_bytes() { return sizeof(DepthRow); }

void launch_depth(hipStream_t st, const char* args_dev, const DepthPlan& plan, const int32_t* vdata, void* rows, int32_t* draw_rows, float* texels)
{
	if(plan.num_draws == 0 || plan.row_blocks == 0)
		return;
	DepthLaunch L{reinterpret_cast<const DepthDrawDev*>(args_dev), reinterpret_cast<const DepthTargetDev*>(args_dev + plan.targets_at),
		plan.num_targets, 0, vdata, static_cast<DepthRow*>(rows), draw_rows, texels};
	hipLaunchKernelGGL(k_depth_setup, dim3(plan.num_draws), dim3(64), 0, st, L);
	hipLaunchKernelGGL(k_depth_raster, dim3(plan.row_blocks), dim3(64), 0, st, L);
}
} // namespace swk
