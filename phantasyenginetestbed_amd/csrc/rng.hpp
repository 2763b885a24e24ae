// PCG32 (O'Neill 2014, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
// Algorithms for Random Number Generation"), identical on host and device. It replaces the
// reference's curand XORWOW states (InitCurandKernel.cu:24-33), whose sequence is not reproduced.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define PHX_HD __host__ __device__ __forceinline__
#else
#define PHX_HD inline
#endif

namespace phx {

struct Pcg32 {
	uint64_t state;
	uint64_t inc;

	PHX_HD uint32_t next()
	{
		uint64_t old = state;
		state = old * 6364136223846793005ull + inc;
		uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
		uint32_t rot = uint32_t(old >> 59u);
		return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
	}

	// Uniform float in [0, 1) with 24 random bits (exact conversion).
	PHX_HD float uniform() { return float(next() >> 8) * (1.0f / 16777216.0f); }

	// pcg32_srandom_r: stream `seq`, start `initState`.
	PHX_HD static Pcg32 seeded(uint64_t initState, uint64_t seq)
	{
		Pcg32 r;
		r.state = 0u;
		r.inc = (seq << 1u) | 1u;
		r.next();
		r.state += initState;
		r.next();
		return r;
	}
};

// Generator of ray `id` in frame `frame` of a run seeded with `seed`.
PHX_HD Pcg32 pixel_rng(uint64_t seed, uint32_t frame, uint32_t id)
{
	return Pcg32::seeded(seed ^ (uint64_t(frame) << 32), uint64_t(id));
}

}  // namespace phx
