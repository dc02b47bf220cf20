// CPU check of the kernels' exact-division helpers (square-wheel_amd/csrc/sw_divide.h), compiled by a
// host compiler from the same header the kernels include. TEST INFRASTRUCTURE ONLY.
//
// Each entry point evaluates one helper on caller-supplied inputs, or sweeps a range itself and
// counts disagreements with the C++ integer operators, which define the expected result.
#include <stdint.h>

#include "../../square-wheel_amd/csrc/sw_divide.h"

extern "C" {

void check_fixed16_div(const int32_t* x, const int32_t* y, int32_t* out, uint64_t n)
{
	for(uint64_t i = 0; i < n; ++i)
		out[i] = swdiv::fixed16_div(x[i], y[i]);
}

void check_div_rcp(const uint32_t* n, const uint32_t* d, uint32_t* q, uint32_t* r, uint64_t count)
{
	for(uint64_t i = 0; i < count; ++i)
		q[i] = swdiv::div_rcp(n[i], d[i], 1.0 / double(d[i]), r[i]);
}

void check_mod_rcp(const uint32_t* n, const uint32_t* d, uint32_t* r, uint64_t count)
{
	for(uint64_t i = 0; i < count; ++i)
		r[i] = swdiv::mod_rcp(n[i], d[i], swdiv::pow2_mask(d[i]), 1.0 / double(d[i]));
}

void check_z_from_inv_z(const uint32_t* div, uint32_t* z, uint64_t count)
{
	for(uint64_t i = 0; i < count; ++i)
		z[i] = swdiv::z_from_inv_z(div[i]);
}

// Every divisor in [d_lo, d_hi] against numerators around 0, 2^32 - 1 and each of the divisor's
// multiples near k * 2^24 (the quotient's range ends and the points where the truncated
// reciprocal product is most likely to be off by one). Returns the number of wrong results.
uint64_t sweep_rcp(uint32_t d_lo, uint32_t d_hi)
{
	uint64_t bad = 0;
	for(uint64_t d = d_lo; d <= d_hi; ++d)
	{
		const uint32_t dd = uint32_t(d);
		const double inv = 1.0 / double(dd);
		const uint32_t mask = swdiv::pow2_mask(dd);
		for(uint64_t k = 0; k < 257; ++k)
		{
			const uint64_t base = k == 256 ? 0xFFFFFFFFull : (k << 24);
			const uint64_t m = (base / d) * d; // a multiple of d at or below base
			for(int64_t off = -2; off <= 2; ++off)
			{
				const int64_t nn = int64_t(m) + off;
				if(nn < 0 || nn > 0xFFFFFFFFll)
					continue;
				const uint32_t n = uint32_t(nn);
				uint32_t rem;
				const uint32_t q = swdiv::div_rcp(n, dd, inv, rem);
				bad += (q != n / dd) || (rem != n % dd);
				bad += swdiv::mod_rcp(n, dd, mask, inv) != n % dd;
			}
		}
	}
	return bad;
}

// fixed16_div for every x in [x_lo, x_hi] and y in a fixed set of divisors, against the i64 quotient.
uint64_t sweep_fixed16_div(int32_t x_lo, int32_t x_hi, const int32_t* ys, uint32_t ny)
{
	uint64_t bad = 0;
	for(int64_t x = x_lo; x <= x_hi; ++x)
		for(uint32_t j = 0; j < ny; ++j)
			if(ys[j] != 0)
				bad += swdiv::fixed16_div(int32_t(x), ys[j]) != int32_t((x * 65536) / int64_t(ys[j]));
	return bad;
}

// The fast branch of z_from_inv_z for every divisor in [lo, hi] (within [1024, 2^24)) with its
// reciprocal off by up to `ulps` units in the last place either way (the device's v_rcp_f32 is
// within one): every result must still be 2^31 / div.
uint64_t sweep_z_rcp_perturbed(uint32_t lo, uint32_t hi, int32_t ulps)
{
	uint64_t bad = 0;
	for(uint64_t d = lo; d <= hi; ++d)
	{
		const float r = 1.0f / float(uint32_t(d));
		uint32_t bits;
		__builtin_memcpy(&bits, &r, 4);
		for(int32_t k = -ulps; k <= ulps; ++k)
		{
			const uint32_t b = bits + uint32_t(k);
			float rk;
			__builtin_memcpy(&rk, &b, 4);
			bad += swdiv::z_from_rcp(uint32_t(d), rk) != uint32_t((1ull << 31) / d);
		}
	}
	return bad;
}

// z_from_inv_z for every divisor in [lo, hi].
uint64_t sweep_z(uint32_t lo, uint32_t hi)
{
	uint64_t bad = 0;
	for(uint64_t d = lo; d <= hi; ++d)
		bad += swdiv::z_from_inv_z(uint32_t(d)) != uint32_t((1ull << 31) / d);
	return bad;
}

} // extern "C"
