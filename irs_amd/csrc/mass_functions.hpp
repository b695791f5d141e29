// Stellar mass functions: inverse-CDF sampling and analytic moments.
// Restates mass_functions/power_law.cuh:30-226 (power laws), kroupa.cuh:22-502
// (three-segment Kroupa), salpeter.cuh:22 (slope -2.35), uniform.cuh (slope 0),
// equal.cuh:27-72 and optical_depth.cuh:28-91. A broken power law is handled
// as the list of its segments that have positive width inside [m_lower,
// m_upper], with the pdf continuous across the breaks; that is the same case
// split the Kroupa class spells out branch by branch.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

namespace mlens
{

enum MassFunctionId
{
	MF_EQUAL = 0,
	MF_UNIFORM = 1,
	MF_SALPETER = 2,
	MF_KROUPA = 3,
	MF_OPTICAL_DEPTH = 4,
	MF_INVALID = -1
};

inline int mass_function_id(const std::string& s)
{
	if (s == "equal") return MF_EQUAL;
	if (s == "uniform") return MF_UNIFORM;
	if (s == "salpeter") return MF_SALPETER;
	if (s == "kroupa") return MF_KROUPA;
	if (s == "optical_depth") return MF_OPTICAL_DEPTH;
	return MF_INVALID;
}

template <typename T>
struct PowerSegments
{
	int n = 0;
	T lo[3], hi[3], slope[3], b[3], prob[3];
};

template <typename T>
__host__ __device__ inline T power_integral(T x1, T x2, T a, T b = 1)
{
	if (a != -1) return b * (std::pow(x2, a + 1) - std::pow(x1, a + 1)) / (a + 1);
	return b * (std::log(x2) - std::log(x1));
}

template <typename T>
__host__ __device__ inline T power_log_integral(T x1, T x2, T a, T b = 1)
{
	if (a != -1)
		return b * (std::pow(x2, a + 1) * ((a + 1) * std::log(x2) - 1) - std::pow(x1, a + 1) * ((a + 1) * std::log(x1) - 1)) / ((a + 1) * (a + 1));
	return b * (std::log(x2) * std::log(x2) - std::log(x1) * std::log(x1)) / 2;
}

template <typename T>
__host__ __device__ inline T invert_power_integral(T p, T x1, T a, T b)
{
	if (a != -1) return std::pow(p * (a + 1) / b + std::pow(x1, a + 1), 1 / (a + 1));
	return x1 * std::exp(p / b);
}

template <typename T>
__host__ __device__ inline PowerSegments<T> power_segments(int mf, T ml, T mu, T ms)
{
	PowerSegments<T> s;
	if (mf == MF_KROUPA)
	{
		const T slopes[3] = {T(-0.3), T(-1.3), T(-2.3)};
		const T m0 = T(0.08) * ms;
		const T m1 = T(0.5) * ms;
		const T lows[3] = {0, m0, m1};
		T k = 1;
		bool first = true;
		T total = 0;
		for (int i = 0; i < 3; i++)
		{
			T lo = ml > lows[i] ? ml : lows[i];
			T hi = (i == 0) ? (mu < m0 ? mu : m0) : (i == 1 ? (mu < m1 ? mu : m1) : mu);
			if (!(lo < hi)) continue;
			if (!first) k *= std::pow(lows[i], slopes[i - 1] - slopes[i]);
			first = false;
			s.lo[s.n] = lo;
			s.hi[s.n] = hi;
			s.slope[s.n] = slopes[i];
			s.b[s.n] = k;
			s.prob[s.n] = power_integral<T>(lo, hi, slopes[i]) * k;
			total += s.prob[s.n];
			s.n++;
		}
		for (int i = 0; i < s.n; i++)
		{
			s.b[i] /= total;
			s.prob[i] /= total;
		}
	}
	else
	{
		s.n = 1;
		s.lo[0] = ml;
		s.hi[0] = mu;
		s.slope[0] = (mf == MF_SALPETER) ? T(-2.35) : T(0);
		s.b[0] = 1 / power_integral<T>(ml, mu, s.slope[0]);
		s.prob[0] = 1;
	}
	return s;
}

// Mass for a uniform draw p in (0, 1].
template <typename T>
__host__ __device__ inline T sample_mass(int mf, T p, T ml, T mu, T ms)
{
	if (mf == MF_EQUAL) return 1;
	if (ml == mu) return ml;
	if (mf == MF_OPTICAL_DEPTH) return (p <= mu / (ml + mu)) ? ml : mu;
	PowerSegments<T> s = power_segments<T>(mf, ml, mu, ms);
	T cum = 0;
	for (int i = 0; i < s.n; i++)
	{
		if (p < cum + s.prob[i] || i == s.n - 1) return invert_power_integral<T>(p - cum, s.lo[i], s.slope[i], s.b[i]);
		cum += s.prob[i];
	}
	return s.lo[0];
}

// <m>, <m^2>, <m^2 ln m>
template <typename T>
inline void mass_moments(int mf, T ml, T mu, T ms, T out[3])
{
	if (mf == MF_EQUAL)
	{
		out[0] = 1;
		out[1] = 1;
		out[2] = 0;
		return;
	}
	if (ml == mu)
	{
		out[0] = ml;
		out[1] = ml * ml;
		out[2] = ml * ml * std::log(ml);
		return;
	}
	if (mf == MF_OPTICAL_DEPTH)
	{
		out[0] = 2 * ml * mu / (ml + mu);
		out[1] = ml * mu;
		out[2] = ml * mu * (ml * std::log(ml) + mu * std::log(mu)) / (ml + mu);
		return;
	}
	PowerSegments<T> s = power_segments<T>(mf, ml, mu, ms);
	out[0] = out[1] = out[2] = 0;
	for (int i = 0; i < s.n; i++)
	{
		out[0] += power_integral<T>(s.lo[i], s.hi[i], s.slope[i] + 1, s.b[i]);
		out[1] += power_integral<T>(s.lo[i], s.hi[i], s.slope[i] + 2, s.b[i]);
		out[2] += power_log_integral<T>(s.lo[i], s.hi[i], s.slope[i] + 2, s.b[i]);
	}
}

}  // namespace mlens
