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

/*
Area of a ccw triangle (a, b, c) inside the unit pixel [0,1]^2, by Green's
theorem: area(T n B) = closed integral of x dy over the boundary of T n B,
which is the part of each edge of T inside B plus the part of B's right edge
(x = 1) inside T. B's left edge has x = 0 and its horizontal edges have dy = 0,
so they contribute nothing. Each edge is clipped to B parametrically
(Y.-D. Liang and B. A. Barsky, ACM TOG 3(1), 1984). No polygon is stored,
so the whole computation stays in registers.
*/
template <typename R>
__device__ __forceinline__ R seg_xdy_in_box(R px, R py, R dx, R dy, R idx, R idy)
{
	R t0 = 0, t1 = 1;
	if (dx != 0)
	{
		R ta = -px * idx, tb = (1 - px) * idx;
		R lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
		t0 = t0 > lo ? t0 : lo;
		t1 = t1 < hi ? t1 : hi;
	}
	else if (px < 0 || px > 1 || (px == 1 && dy > 0))
	{
		// a vertical edge on x = 1 with the triangle on its left is the part
		// of the pixel's right edge inside the triangle, which the x = 1 term
		// of tri_pixel_area already counts
		return 0;
	}
	if (dy != 0)
	{
		R ta = -py * idy, tb = (1 - py) * idy;
		R lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
		t0 = t0 > lo ? t0 : lo;
		t1 = t1 < hi ? t1 : hi;
	}
	else if (py < 0 || py > 1)
	{
		return 0;
	}
	if (!(t0 < t1)) return 0;
	R xa = px + t0 * dx, xb = px + t1 * dx;
	return (xa + xb) * R(0.5) * ((t1 - t0) * dy);
}

template <typename R>
__device__ __forceinline__ void vline_hit(R px, R py, R dx, R dy, R idx, R X, R& lo, R& hi)
{
	R qx = px + dx;
	if (dx != 0)
	{
		if ((px - X) * (qx - X) <= 0)
		{
			R y = py + (X - px) * idx * dy;
			lo = lo < y ? lo : y;
			hi = hi > y ? hi : y;
		}
	}
	else if (px == X)
	{
		R y2 = py + dy;
		R l = py < y2 ? py : y2, h = py < y2 ? y2 : py;
		lo = lo < l ? lo : l;
		hi = hi > h ? hi : h;
	}
}

template <typename R>
struct Tri
{
	R x[3], y[3];      // ccw vertices, relative to an integer origin
	R dx[3], dy[3];    // edge vectors v[k+1] - v[k]
	R idx[3], idy[3];  // their reciprocals (unused when the component is zero)
};

template <typename R>
__device__ __forceinline__ R recip(R v);
template <>
__device__ __forceinline__ float recip<float>(float v) { return rcp32(v); }
template <>
__device__ __forceinline__ double recip<double>(double v) { return rcp64(v); }

template <typename R>
__device__ __forceinline__ void tri_setup(Tri<R>& t)
{
#pragma unroll
	for (int k = 0; k < 3; k++)
	{
		int n = (k + 1) % 3;
		t.dx[k] = t.x[n] - t.x[k];
		t.dy[k] = t.y[n] - t.y[k];
		t.idx[k] = t.dx[k] != 0 ? recip<R>(t.dx[k]) : R(0);
		t.idy[k] = t.dy[k] != 0 ? recip<R>(t.dy[k]) : R(0);
	}
}

// area of the triangle inside pixel [X, X+1] x [Y, Y+1] (X, Y relative to the triangle origin)
template <typename R>
__device__ __forceinline__ R tri_pixel_area(const Tri<R>& t, R X, R Y)
{
	R a = 0;
	// y-range of the triangle on the line x = 1, clipped to [0, 1] below, so
	// an empty range can start as [1, 0]
	R lo = R(1), hi = R(0);
#pragma unroll
	for (int k = 0; k < 3; k++)
	{
		R px = t.x[k] - X, py = t.y[k] - Y;
		a += seg_xdy_in_box<R>(px, py, t.dx[k], t.dy[k], t.idx[k], t.idy[k]);
		vline_hit<R>(px, py, t.dx[k], t.dy[k], t.idx[k], R(1), lo, hi);
	}
	lo = lo > 0 ? lo : R(0);
	hi = hi < 1 ? hi : R(1);
	if (hi > lo) a += hi - lo;
	return a;
}

// y-range of the triangle inside the vertical strip [xl, xr]
template <typename R>
__device__ __forceinline__ void tri_strip_yrange(const Tri<R>& t, R xl, R xr, R& ylo, R& yhi)
{
	ylo = R(1e30);
	yhi = R(-1e30);
#pragma unroll
	for (int k = 0; k < 3; k++)
	{
		R px = t.x[k], py = t.y[k], dx = t.dx[k], dy = t.dy[k];
		R t0 = 0, t1 = 1;
		if (dx != 0)
		{
			R ta = (xl - px) * t.idx[k], tb = (xr - px) * t.idx[k];
			R lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
			t0 = t0 > lo ? t0 : lo;
			t1 = t1 < hi ? t1 : hi;
		}
		else if (px < xl || px > xr)
		{
			continue;
		}
		if (t0 > t1) continue;
		R ya = py + t0 * dy, yb = py + t1 * dy;
		R l = ya < yb ? ya : yb, h = ya < yb ? yb : ya;
		ylo = ylo < l ? ylo : l;
		yhi = yhi > h ? yhi : h;
	}
}

}  // namespace dev
}  // namespace mlens
