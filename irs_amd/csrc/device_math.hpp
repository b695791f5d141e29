// Device-side arithmetic shared by the FMM and shooting kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mlens
{
namespace dev
{

struct dcx
{
	double re, im;
};

__device__ __forceinline__ dcx mk(double r, double i) { return dcx{r, i}; }
__device__ __forceinline__ dcx add(dcx a, dcx b) { return dcx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ dcx sub(dcx a, dcx b) { return dcx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ dcx mul(dcx a, dcx b) { return dcx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ dcx scale(dcx a, double s) { return dcx{a.re * s, a.im * s}; }
__device__ __forceinline__ dcx conj(dcx a) { return dcx{a.re, -a.im}; }
__device__ __forceinline__ double cabs(dcx a) { return hypot(a.re, a.im); }
__device__ __forceinline__ dcx clog(dcx a) { return dcx{log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }
__device__ __forceinline__ dcx cexp(dcx a) { return dcx{exp(a.re) * cos(a.im), exp(a.re) * sin(a.im)}; }

// Smith's complex division (R. L. Smith, CACM 5(8), 1962), as complex.cuh:247-262 uses.
__device__ __forceinline__ dcx cdiv(dcx a, dcx b)
{
	if (fabs(b.im) < fabs(b.re))
	{
		double f = b.im / b.re;
		double d = b.re + b.im * f;
		return dcx{(a.re + a.im * f) / d, (a.im - a.re * f) / d};
	}
	double f = b.re / b.im;
	double d = b.re * f + b.im;
	return dcx{(a.re * f + a.im) / d, (a.im * f - a.re) / d};
}

__device__ __forceinline__ dcx cpow(dcx a, int n)
{
	dcx r = dcx{1, 0};
	for (int i = 0; i < n; i++) r = mul(r, a);
	return r;
}

// 1/x in FP64: hardware estimate refined by two Newton-Raphson steps (quadratic
// convergence takes the estimate to full double precision).
__device__ __forceinline__ double rcp64(double x)
{
	double y = __builtin_amdgcn_rcp(x);
	double e = fma(-x, y, 1.0);
	y = fma(y, e, y);
	e = fma(-x, y, 1.0);
	y = fma(y, e, y);
	return y;
}

__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

/* Smooth-matter deflection, alpha_smooth.cuh:62-127, in FP64. */
__device__ inline dcx alpha_smooth(dcx z, double ks, int rectangular, int approx, int taylor_smooth, dcx corner)
{
	const double inv_pi = 0.31830988618379067154;
	dcx a;
	if (rectangular)
	{
		dcx zb = conj(z);
		if (approx)
		{
			dcx s1 = {0, 0}, s2 = {0, 0}, s3 = {0, 0}, s4 = {0, 0};
			dcx r1 = cdiv(zb, corner), r2 = cdiv(zb, conj(corner));
			dcx r3 = cdiv(zb, dcx{-corner.re, -corner.im}), r4 = cdiv(zb, dcx{-corner.re, corner.im});
			for (int i = taylor_smooth; i >= 1; i--)
			{
				double inv = 1.0 / i;
				s1.re += inv;
				s2.re += inv;
				s3.re += inv;
				s4.re += inv;
				s1 = mul(s1, r1);
				s2 = mul(s2, r2);
				s3 = mul(s3, r3);
				s4 = mul(s4, r4);
			}
			dcx c = corner, cb = conj(corner), mc = dcx{-corner.re, -corner.im}, mcb = dcx{-corner.re, corner.im};
			a = sub(add(sub(mul(sub(c, zb), sub(clog(c), s1)), mul(sub(cb, zb), sub(clog(cb), s2))),
						mul(sub(mc, zb), sub(clog(mc), s3))),
				mul(sub(mcb, zb), sub(clog(mcb), s4)));
			a = mul(a, dcx{0, -ks * inv_pi});
			a.re -= ks * 2 * (corner.re + z.re);
		}
		else
		{
			dcx c1 = sub(corner, zb), c2 = sub(conj(corner), zb);
			dcx c3 = sub(dcx{-corner.re, -corner.im}, zb), c4 = sub(dcx{-corner.re, corner.im}, zb);
			a = sub(add(sub(mul(c1, clog(c1)), mul(c2, clog(c2))), mul(c3, clog(c3))), mul(c4, clog(c4)));
			a = mul(a, dcx{0, -ks * inv_pi});
			double box = (-corner.re <= z.re && z.re <= corner.re && -corner.im <= z.im && z.im <= corner.im) ? 1 : 0;
			a.re -= ks * 2 * (corner.re + z.re) * box;
			double h = (corner.im + z.im >= 0 && corner.im - z.im >= 0 && z.re - corner.re >= 0) ? 1 : 0;
			a.re -= ks * 4 * corner.re * h;
		}
	}
	else
	{
		double cabs_ = hypot(corner.re, corner.im);
		if (approx || hypot(z.re, z.im) <= cabs_)
			a = dcx{-ks * z.re, -ks * z.im};
		else
			a = cdiv(dcx{-ks * cabs_ * cabs_, 0}, conj(z));
	}
	return a;
}

/* alpha_smooth.cuh:141-236: d alpha_smooth / dz (real) and d alpha_smooth / dzbar */
__device__ inline void d_alpha_smooth(dcx z, double ks, int rectangular, int approx, int taylor_smooth, dcx corner,
	double& ddz, dcx& ddzb)
{
	const double inv_pi = 0.31830988618379067154;
	double box = (-corner.re <= z.re && z.re <= corner.re && -corner.im <= z.im && z.im <= corner.im) ? 1 : 0;
	if (rectangular)
	{
		if (approx)
		{
			ddz = -ks;
			dcx r = cdiv(conj(z), corner);
			double ph = atan2(corner.im, corner.re);
			dcx d = {0, 0};
			for (int i = taylor_smooth - 1; i >= 2; i -= 2)
			{
				dcx e = cexp(dcx{0, 2 * ph * i});
				d = add(d, dcx{(1 - e.re) / i, -e.im / i});
				d = mul(d, mul(r, r));
			}
			d = scale(d, 2);
			d = mul(d, dcx{0, -ks * inv_pi});
			d.re += ks - 4 * ks * ph * inv_pi;
			ddzb = d;
		}
		else
		{
			ddz = -ks * box;
			dcx zb = conj(z);
			dcx c1 = sub(conj(corner), zb), c2 = sub(corner, zb);
			dcx c3 = sub(dcx{-corner.re, -corner.im}, zb), c4 = sub(dcx{-corner.re, corner.im}, zb);
			dcx d = add(sub(sub(clog(c1), clog(c2)), clog(c3)), clog(c4));
			d = mul(d, dcx{0, -ks * inv_pi});
			d.re -= ks * box;
			ddzb = d;
		}
	}
	else
	{
		double cabs_ = hypot(corner.re, corner.im);
		if (approx)
		{
			ddz = -ks;
			ddzb = dcx{0, 0};
		}
		else
		{
			ddz = -ks * ((cabs_ - hypot(z.re, z.im) >= 0) ? 1 : 0);
			dcx zb = conj(z);
			ddzb = cdiv(dcx{ks * cabs_ * cabs_, 0}, mul(zb, zb));
		}
	}
}

/* alpha_smooth.cuh:252-316: d^2 alpha_smooth / dzbar^2 */
__device__ inline dcx d2_alpha_smooth(dcx z, double ks, int rectangular, int approx, int taylor_smooth, dcx corner)
{
	const double inv_pi = 0.31830988618379067154;
	dcx zb = conj(z);
	if (rectangular)
	{
		if (approx)
		{
			// the derivative of the even series of d alpha_smooth / dzbar, term by term
			dcx r = cdiv(zb, corner);
			double ph = atan2(corner.im, corner.re);
			dcx d = {0, 0};
			for (int i = taylor_smooth - 1; i >= 2; i -= 2)
			{
				dcx e = cexp(dcx{0, 2 * ph * i});
				d = add(d, dcx{1 - e.re, -e.im});
				d = mul(d, mul(r, r));
			}
			d = cdiv(d, zb);
			d = scale(d, 2);
			return mul(d, dcx{0, -ks * inv_pi});
		}
		dcx c1 = sub(conj(corner), zb), c2 = sub(corner, zb);
		dcx c3 = sub(dcx{-corner.re, -corner.im}, zb), c4 = sub(dcx{-corner.re, corner.im}, zb);
		dcx one = {1, 0};
		dcx d = sub(add(add(scale(cdiv(one, c1), -1), cdiv(one, c2)), cdiv(one, c3)), cdiv(one, c4));
		return mul(d, dcx{0, -ks * inv_pi});
	}
	if (approx) return dcx{0, 0};
	double cabs_ = hypot(corner.re, corner.im);
	return cdiv(dcx{-2 * ks * cabs_ * cabs_, 0}, mul(zb, mul(zb, zb)));
}

/*
Area of a counter-clockwise triangle inside the pixel [X, X+1] x [Y, Y+1], by
Green's theorem with Q(x, y) = clamp(x - X, 0, 1) on the row Y <= y <= Y + 1
and 0 elsewhere: dQ/dx is the pixel's indicator, so the area is the sum over
the edges of the integral of Q dy. On the part of an edge inside the row, with
u = x - X linear in y between ua and ub, the mean of clamp(u, 0, 1) is
  ((H - L)(H + L) / 2 + max(hi, 1) - max(lo, 1)) / (hi - lo),
where lo, hi are ua, ub in order and L, H their values clamped to [0, 1]
(clamp(lo, 0, 1) when lo = hi). Both ratios stay in [0, 1], so a near-vertical
edge loses no accuracy; there are no branches.
k[i] = dx_i / dy_i of edge i (0 for a horizontal edge, whose part inside the
row has no height).
*/
__device__ __forceinline__ double tri_pixel_area(const double x[3], const double y[3], const double k[3], double X, double Y)
{
	double a = 0;
#pragma unroll
	for (int i = 0; i < 3; i++)
	{
		const int n = (i + 1) % 3;
		const double v0 = y[i] - Y, v1 = y[n] - Y, u0 = x[i] - X;
		const double va = fmin(fmax(v0, 0.0), 1.0), vb = fmin(fmax(v1, 0.0), 1.0);
		const double ua = fma(va - v0, k[i], u0), ub = fma(vb - v0, k[i], u0);
		const double lo = fmin(ua, ub), hi = fmax(ua, ub);
		const double L = fmin(fmax(lo, 0.0), 1.0), H = fmin(fmax(hi, 0.0), 1.0);
		const double d = hi - lo;
		const double num = fma((H - L) * 0.5, H + L, fmax(hi, 1.0) - fmax(lo, 1.0));
		const double mean = d > 0 ? num * rcp64(d) : L;
		a = fma(vb - va, mean, a);
	}
	return a;
}

}  // namespace dev
}  // namespace mlens
