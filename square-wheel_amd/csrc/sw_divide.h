// Exact integer divisions computed through double-precision arithmetic.
//
// The kernels divide 64-bit and 32-bit integers in their inner loops; gfx950 has no integer divide
// instruction, and the compiler's software division costs tens of instructions. Each helper here
// takes the correctly rounded double quotient (or the product with a double reciprocal), truncates
// it and corrects it by one step where needed, which is exact for the whole input range stated.
//
// This header carries no HIP dependency: the kernels include it with SW_HD defined as
// __host__ __device__ __forceinline__ (sw_device.h), and a CPU check compiles the same functions
// with a host compiler to compare them with integer division over random and edge inputs.
#pragma once

#include <stdint.h>

#ifndef SW_HD
#define SW_HD static inline
#endif

namespace swdiv
{

const uint32_t NO_MASK = 0xFFFFFFFFu;

// ((x as i64) << 16) / y, truncated toward zero (fixed_math.rs:56-59). |x << 16| < 2^47 is exact in
// a double; the correctly rounded double quotient, truncated, is off by at most one, which one
// integer step corrects. y = 0 (the reference would panic) gives 0.
SW_HD int32_t fixed16_div(int32_t x, int32_t y)
{
	if(y == 0)
		return 0;
	const int64_t a = int64_t(x) * 65536;
	const uint64_t ua = a < 0 ? uint64_t(-a) : uint64_t(a);
	const uint64_t ub = y < 0 ? uint64_t(-int64_t(y)) : uint64_t(y);
	uint64_t q = uint64_t(double(ua) / double(ub));
	const int64_t r = int64_t(ua) - int64_t(q * ub);
	if(r < 0)
		q -= 1;
	else if(r >= int64_t(ub))
		q += 1;
	return int32_t((a < 0) != (y < 0) ? -int64_t(q) : int64_t(q));
}

// n mod d for 1 <= d <= 2^30: a power of two by its mask, otherwise through the double reciprocal
// inv = 1 / d, whose truncated product with n is within one of n / d for every 32-bit n (relative
// error below 2^-51, quotient below 2^32), so one correction either way makes it exact.
// (The corrections are written as selects: per-lane branches would cost the kernels execution-mask
// changes around two-instruction blocks.)
SW_HD uint32_t mod_rcp(uint32_t n, uint32_t d, uint32_t mask, double inv)
{
	if(mask != NO_MASK)
		return n & mask;
	const uint32_t q = uint32_t(double(n) * inv);
	const int32_t r = int32_t(n - q * d);
	const int32_t rn = r < 0 ? r + int32_t(d) : r;
	return uint32_t(rn >= int32_t(d) ? rn - int32_t(d) : rn);
}

// n / d and n mod d for 1 <= d <= 2^30, exact as mod_rcp.
SW_HD uint32_t div_rcp(uint32_t n, uint32_t d, double inv, uint32_t& rem)
{
	const uint32_t q = uint32_t(double(n) * inv);
	const int32_t r = int32_t(n - q * d);
	const bool lo = r < 0, hi = r >= int32_t(d);
	rem = uint32_t(lo ? r + int32_t(d) : hi ? r - int32_t(d) : r);
	return lo ? q - 1u : hi ? q + 1u : q;
}

SW_HD uint32_t pow2_mask(uint32_t d) { return (d & (d - 1u)) == 0u ? d - 1u : NO_MASK; }

// Approximate f32 reciprocal: the device's v_rcp_f32 (at most 1 ulp from 1 / x), the host's correctly
// rounded 1 / x.
#if defined(__HIP_DEVICE_COMPILE__)
#define SWDIV_RCP(x) __builtin_amdgcn_rcpf(x)
#else
#define SWDIV_RCP(x) (1.0f / (x))
#endif

// 2^31 / div for div >= 1 (the perspective span's z, rasterizer.rs:430), exact.
//   1024 <= div < 2^24: div is exact in f32 and the quotient x = 2^31 / div is at most 2^21, so
//   2^31 * rcp(div) is within x * 2^-22 <= 0.5 of x (an rcp within 2 ulp would still do): its
//   truncation q is floor(x) - 1, floor(x) or floor(x) + 1, and the remainder 2^31 - q * div (no
//   u32 overflow: q * div <= 2^31 + div / 2) picks floor(x) with one step.
//   Otherwise: the correctly rounded double quotient truncates to the integer quotient exactly,
//   since the exact quotient's fraction is 0 or between 1 / div and 1 - 1 / div, while the rounding
//   error is at most 2^-53 * 2^31 / div.
SW_HD uint32_t z_from_rcp(uint32_t div, float rcp)
{
	uint32_t q = uint32_t(rcp * 2147483648.0f);
	const int32_t rem = int32_t(0x80000000u - q * div);
	if(rem < 0)
		q -= 1u;
	else if(rem >= int32_t(div))
		q += 1u;
	return q;
}

SW_HD uint32_t z_from_inv_z(uint32_t div)
{
	if(div >= 1024u && div < (1u << 24))
		return z_from_rcp(div, SWDIV_RCP(float(div)));
	return uint32_t(double(1u << 31) / double(div));
}

} // namespace swdiv
