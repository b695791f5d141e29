// Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes, R. O. Dror,
// D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), with the
// published multipliers and Weyl increments. The star field replaces the
// reference's curand XORWOW stream (star.cuh:40-107): star i draws its uniforms
// from counters (i, 0) and (i, 1) under key (seed, 0), so a field is a pure
// function of (seed, i) on any device or host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mlens
{

__host__ __device__ inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
	uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
	uint32_t k0 = key_in[0], k1 = key_in[1];
#pragma unroll
	for (int r = 0; r < 10; r++)
	{
		if (r > 0)
		{
			k0 += 0x9E3779B9u;
			k1 += 0xBB67AE85u;
		}
		uint64_t p0 = uint64_t(0xD2511F53u) * c0;
		uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
		uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
		uint32_t n1 = uint32_t(p1);
		uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
		uint32_t n3 = uint32_t(p0);
		c0 = n0;
		c1 = n1;
		c2 = n2;
		c3 = n3;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

// 53 random bits from two words, mapped onto (0, 1] like curand_uniform_double.
__host__ __device__ inline double uniform_open_closed(uint32_t hi, uint32_t lo)
{
	uint64_t x = ((uint64_t(hi) << 32) | lo) >> 11;
	return double(x + 1) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline void star_uniforms(uint32_t seed, long long i, double u[4])
{
	uint32_t key[2] = {seed, 0u};
	uint32_t ctr[4] = {uint32_t(i), uint32_t(uint64_t(i) >> 32), 0u, 0u};
	uint32_t o[4];
	philox4x32_10(ctr, key, o);
	u[0] = uniform_open_closed(o[0], o[1]);
	u[1] = uniform_open_closed(o[2], o[3]);
	ctr[2] = 1u;
	philox4x32_10(ctr, key, o);
	u[2] = uniform_open_closed(o[0], o[1]);
	u[3] = uniform_open_closed(o[2], o[3]);
}

}  // namespace mlens
