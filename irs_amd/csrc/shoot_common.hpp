// Shared by the shooting kernels and the critical-curve solver: the kernel
// argument block, the tree descent, the lens equation at a leaf or at any
// tree node, and the host-side builder of the argument block.
#pragma once

#include "engine.hpp"
#include "device_math.hpp"

#include <cstring>
#include <stdexcept>

namespace mlens
{

using dev::dcx;

static inline unsigned blocks_for(long long n, int threads) { return unsigned((n + threads - 1) / threads); }

struct ShootArgs
{
	// lens
	double kappa, gamma, theta2, kappastar;
	int rectangular, approx, taylor_smooth, smooth_linear;
	double corner_re, corner_im;
	double lin_x, lin_y;   // linear part of the lens map (with -alpha_smooth folded in when it is linear)
	// lattice: point (I, J) = (x0 + stepx I, y0 + stepy J)
	double x0, y0, stepx, stepy;
	long long N1, N2;
	// source-plane map
	double cyx, cyy, h1, h2, s1, s2, factor;
	int n1, n2;
	int tiles_x;   // tiles per row of a tiled map (map_index)
	// tree
	const double *ncx, *ncy, *nhl;
	const int *nchild0, *nbegin, *ncount, *nnnbr, *nnbr, *leaf_of_node;
	const double* local;
	int p;
	const double *sx, *sy, *sm;
	// leaf tables
	const double* leaf_geom;
	const int* leaf_near_n;
	const double* near_d;
	const float* near_f;
	const double* poly_d;
	const float* poly_f;
	// maps
	void *map, *mmin, *msad;
	// the fast IPM kernel's fully covered row runs, as differences along each
	// row (tiled maps of the map's type): span (the map, or minima with
	// parities), span_sad
	void *span, *span_sad;
};

// The hot read-only tables travel as separate __restrict__ kernel parameters:
// the compiler then knows that the map atomics cannot modify them, so their
// wave-uniform reads become scalar (SMEM) loads through the scalar cache.
#define MLENS_TABLE_PARAMS                                                                              \
	const double* __restrict__ ncx, const double* __restrict__ ncy, const double* __restrict__ nhl,      \
		const int* __restrict__ nchild0, const int* __restrict__ leaf_of_node,                           \
		const double* __restrict__ leaf_geom, const int* __restrict__ leaf_near_n,                       \
		const double* __restrict__ near_d, const float* __restrict__ near_f,                             \
		const double* __restrict__ poly_d, const float* __restrict__ poly_f
#define MLENS_TABLE_ARGS(a) \
	a.ncx, a.ncy, a.nhl, a.nchild0, a.leaf_of_node, a.leaf_geom, a.leaf_near_n, a.near_d, a.near_f, a.poly_d, a.poly_f
#define MLENS_TABLE_REBIND(a)   \
	a.ncx = ncx;                \
	a.ncy = ncy;                \
	a.nhl = nhl;                \
	a.nchild0 = nchild0;        \
	a.leaf_of_node = leaf_of_node; \
	a.leaf_geom = leaf_geom;    \
	a.leaf_near_n = leaf_near_n; \
	a.near_d = near_d;          \
	a.near_f = near_f;          \
	a.poly_d = poly_d;          \
	a.poly_f = poly_f;

/* Lattice coordinate x0 + step * i with two roundings, never one fused
   multiply-add. Lattice points land exactly on the tree's dividing lines (a
   field centred at the origin puts corners on x = 0 and y = 0), where the
   descent's >= rule picks a side; the FMM deflection jumps across a leaf
   boundary by about its truncation error, so the point must fall on the side
   the host (which has no FMA) puts it. */
__device__ __forceinline__ double lattice_at(double x0, double step, long long i)
{
#pragma clang fp contract(off)
	const double t = step * double(i);
	return x0 + t;
}

__device__ __forceinline__ int quadrant(double x, double y, double cx, double cy)
{
	if (y >= cy) return (x >= cx) ? 0 : 1;
	return (x < cx) ? 2 : 3;
}

// tree_node.cuh:529-561
__device__ __forceinline__ int find_node(const ShootArgs& a, double x, double y)
{
	int node = 0;
	while (true)
	{
		int c0 = a.nchild0[node];
		if (c0 < 0) return node;
		double cx = a.ncx[node], cy = a.ncy[node], h = a.nhl[node];
		if (fabs(x - cx) > h || fabs(y - cy) > h) return node;
		node = c0 + quadrant(x, y, cx, cy);
	}
}

struct Deflection
{
	double wx, wy, mu;
};

/* lens equation at a leaf (leaf index L is wave-uniform) */
template <int PREC, bool MU>
__device__ __forceinline__ Deflection deflect_leaf(const ShootArgs& a, int L, double zx, double zy)
{
	const double* g = a.leaf_geom + 4 * (long long)L;
	const double cx = g[0], cy = g[1], ih = g[3];
	const int packed = a.leaf_near_n[L];   // near stars (padded to 4) | FP32 entries to skip << 16
	const int nn = packed & 0xffff;
	const int p = a.p;
	const double dzx = (zx - cx) * ih, dzy = (zy - cy) * ih;
	double ax = 0, ay = 0, px = 0, py = 0;
	if (PREC == PREC_FP64 || MU)
	{
		const double* s = a.near_d + (long long)L * MAX_NUM_STARS_DIRECT * 3;
		for (int k0 = 0; k0 < nn; k0 += 4)
		{
#pragma unroll
			for (int k = k0; k < k0 + 4; k++)
			{
				double dx = zx - s[3 * k], dy = zy - s[3 * k + 1];
				double f = s[3 * k + 2] * dev::rcp64(dx * dx + dy * dy);
				ax = fma(f, dx, ax);
				ay = fma(f, dy, ay);
			}
		}
		const int SD = poly_stride_d(p);
		const double* A = a.poly_d + (long long)L * SD * 2;
		for (int j0 = 0; j0 < SD; j0 += 4)
		{
#pragma unroll
			for (int j = j0; j < j0 + 4; j++)
			{
				double nx = fma(px, dzx, fma(-py, dzy, A[2 * j]));
				double ny = fma(px, dzy, fma(py, dzx, A[2 * j + 1]));
				px = nx;
				py = ny;
			}
		}
	}
	else
	{
		// FP32 in leaf-relative coordinates; the leaf-constant term A_1 stays FP64
		const float zrx = float(zx - cx), zry = float(zy - cy);
		const float4* s = reinterpret_cast<const float4*>(a.near_f) + (long long)L * MAX_NUM_STARS_DIRECT;
		// The leaf's tables are read with scalar loads, and a wave waits for each batch before using it.
		// Eight entries per trip instead of four halve those waits; the remainder of four (lists and
		// polynomials are padded to fours) keeps the order of summation. Deflection alone on the
		// 10000^2 map: 0.632 -> 0.610 s.
		float fx = 0, fy = 0;
		auto star = [&](int k) {
			const float4 v = s[k];
			const float dx = zrx - v.x, dy = zry - v.y;
			const float f = v.z * dev::rcp32(dx * dx + dy * dy);
			fx = fmaf(f, dx, fx);
			fy = fmaf(f, dy, fy);
		};
		int k0 = 0;
		for (; k0 + 8 <= nn; k0 += 8)
		{
#pragma unroll
			for (int k = k0; k < k0 + 8; k++) star(k);
		}
		if (k0 < nn)
		{
#pragma unroll
			for (int k = k0; k < k0 + 4; k++) star(k);
		}
		ax = fx;
		ay = fy;
		float qx = 0, qy = 0;
		const int SF = poly_stride_f(p);
		const float* A = a.poly_f + (long long)L * SF * 2;
		const float fdx = float(dzx), fdy = float(dzy);
		auto term = [&](int j) {
			const float nx = fmaf(qx, fdx, fmaf(-qy, fdy, A[2 * j]));
			const float ny = fmaf(qx, fdy, fmaf(qy, fdx, A[2 * j + 1]));
			qx = nx;
			qy = ny;
		};
		int j0 = packed >> 16;
		if ((SF - j0) & 4)
		{
#pragma unroll
			for (int j = j0; j < j0 + 4; j++) term(j);
			j0 += 4;
		}
		for (; j0 < SF; j0 += 8)
		{
#pragma unroll
			for (int j = j0; j < j0 + 8; j++) term(j);
		}
		if (p >= 1)
		{
			const int SD = poly_stride_d(p);
			const double* A1 = a.poly_d + ((long long)L * SD + SD - 1) * 2;   // A_1
			px = A1[0] + (dzx * double(qx) - dzy * double(qy));
			py = A1[1] + (dzx * double(qy) + dzy * double(qx));
		}
	}
	Deflection d;
	d.wx = a.lin_x * zx - a.theta2 * ax - px;
	d.wy = a.lin_y * zy - a.theta2 * ay + py;
	d.mu = 0;
	if (!a.smooth_linear)
	{
		dcx sm = dev::alpha_smooth(dcx{zx, zy}, a.kappastar, a.rectangular, a.approx, a.taylor_smooth, dcx{a.corner_re, a.corner_im});
		d.wx -= sm.re;
		d.wy -= sm.im;
	}
	if (MU)
	{
		// ipm_functions.cuh:67-82 with alpha_star.cuh:54-77 and alpha_local.cuh:51-69
		const double* s = a.near_d + (long long)L * MAX_NUM_STARS_DIRECT * 3;
		double sx = 0, sy = 0;   // sum m / (z - z_i)^2
		for (int k = 0; k < nn; k++)
		{
			double dx = zx - s[3 * k], dy = zy - s[3 * k + 1];
			double r2 = dx * dx + dy * dy;
			double ir4 = s[3 * k + 2] == 0 ? 0.0 : 1.0 / (r2 * r2);
			double m = s[3 * k + 2];
			sx += m * (dx * dx - dy * dy) * ir4;
			sy += m * (-2 * dx * dy) * ir4;
		}
		// d alpha_star / d zbar = conj(-theta^2 * S)
		double dsx = -a.theta2 * sx, dsy = a.theta2 * sy;
		// d alpha_local / d zbar = conj( (1/h) dP/d(dz) ), P(dz) = sum_{i>=1} A_i dz^(i-1)
		double lx = 0, ly = 0;
		{
			const int SD = poly_stride_d(p);
			const double* A = a.poly_d + (long long)L * SD * 2;
			for (int j = 0; j < SD - 1; j++)
			{
				double deg = double(SD - 1 - j);
				double nx = lx * dzx - ly * dzy + deg * A[2 * j];
				double ny = lx * dzy + ly * dzx + deg * A[2 * j + 1];
				lx = nx;
				ly = ny;
			}
			lx *= ih;
			ly *= -ih;
		}
		double ddz;
		dcx ddzb;
		dev::d_alpha_smooth(dcx{zx, zy}, a.kappastar, a.rectangular, a.approx, a.taylor_smooth, dcx{a.corner_re, a.corner_im}, ddz, ddzb);
		double dwdz = (1 - a.kappa) - ddz;
		double bx = a.gamma - dsx - lx - ddzb.re;
		double by = -dsy - ly - ddzb.im;
		double muinv = dwdz * dwdz - (bx * bx + by * by);
		d.mu = 1 / muinv;
	}
	return d;
}

/* lens equation at a node that is not a leaf (a point outside the root box,
   tree_node.cuh:535-540): the node's own stars and neighbours' stars directly,
   plus its local expansion, all in FP64 */
template <bool MU>
__device__ Deflection deflect_node(const ShootArgs& a, int node, double zx, double zy)
{
	double ax = 0, ay = 0, sx = 0, sy = 0;
	auto range = [&](int b, int n) {
		for (int s = b; s < b + n; s++)
		{
			double dx = zx - a.sx[s], dy = zy - a.sy[s];
			double r2 = dx * dx + dy * dy;
			double f = a.sm[s] / r2;
			ax += f * dx;
			ay += f * dy;
			if (MU)
			{
				double ir4 = 1.0 / (r2 * r2);
				sx += a.sm[s] * (dx * dx - dy * dy) * ir4;
				sy += a.sm[s] * (-2 * dx * dy) * ir4;
			}
		}
	};
	range(a.nbegin[node], a.ncount[node]);
	for (int j = 0; j < a.nnnbr[node]; j++)
	{
		int nb = a.nnbr[8 * node + j];
		range(a.nbegin[nb], a.ncount[nb]);
	}
	const double h = a.nhl[node];
	const double dzx = (zx - a.ncx[node]) / h, dzy = (zy - a.ncy[node]) / h;
	double px = 0, py = 0, lx = 0, ly = 0;
	const double* c = a.local + 2 * (long long)node * (a.p + 1);
	for (int i = a.p; i >= 1; i--)
	{
		double nx = px * dzx - py * dzy + i * c[2 * i];
		double ny = px * dzy + py * dzx + i * c[2 * i + 1];
		px = nx;
		py = ny;
	}
	px *= a.theta2 / h;
	py *= a.theta2 / h;
	Deflection d;
	d.wx = a.lin_x * zx - a.theta2 * ax - px;
	d.wy = a.lin_y * zy - a.theta2 * ay + py;
	d.mu = 0;
	if (!a.smooth_linear)
	{
		dcx s = dev::alpha_smooth(dcx{zx, zy}, a.kappastar, a.rectangular, a.approx, a.taylor_smooth, dcx{a.corner_re, a.corner_im});
		d.wx -= s.re;
		d.wy -= s.im;
	}
	if (MU)
	{
		for (int i = a.p; i >= 2; i--)
		{
			double nx = lx * dzx - ly * dzy + double(i) * (i - 1) * c[2 * i];
			double ny = lx * dzy + ly * dzx + double(i) * (i - 1) * c[2 * i + 1];
			lx = nx;
			ly = ny;
		}
		lx *= a.theta2 / (h * h);
		ly *= -a.theta2 / (h * h);
		double dsx = -a.theta2 * sx, dsy = a.theta2 * sy;
		double ddz;
		dcx ddzb;
		dev::d_alpha_smooth(dcx{zx, zy}, a.kappastar, a.rectangular, a.approx, a.taylor_smooth, dcx{a.corner_re, a.corner_im}, ddz, ddzb);
		double dwdz = (1 - a.kappa) - ddz;
		double bx = a.gamma - dsx - lx - ddzb.re;
		double by = -dsy - ly - ddzb.im;
		d.mu = 1 / (dwdz * dwdz - (bx * bx + by * by));
	}
	return d;
}

/* ---------------------------------------------------------------------------
Host side
--------------------------------------------------------------------------- */
struct ShootArgsBuilder
{
	static ShootArgs make(const LensParams& L, const GridParams& g, bool irs_offset)
	{
		ShootArgs a;
		std::memset(&a, 0, sizeof(a));
		a.kappa = L.kappa;
		a.gamma = L.gamma;
		a.theta2 = L.theta * L.theta;
		a.kappastar = L.kappastar;
		a.rectangular = L.rectangular;
		a.approx = L.approx;
		a.taylor_smooth = L.taylor_smooth;
		a.corner_re = L.corner_re;
		a.corner_im = L.corner_im;
		a.smooth_linear = (!L.rectangular && L.approx) ? 1 : 0;
		double fold = a.smooth_linear ? L.kappastar : 0.0;   // -alpha_smooth = +kappa_star z (alpha_smooth.cuh:109-112)
		a.lin_x = (1 - L.kappa) + L.gamma + fold;
		a.lin_y = (1 - L.kappa) - L.gamma + fold;
		// lattice origin: corners at (c - h) + 2 rhs I; IRS rays at ((c - h) + rhs) + 2 rhs i (irs_functions.cuh:122)
		a.x0 = g.center_x[0] - g.half_length_x[0];
		a.y0 = g.center_x[1] - g.half_length_x[1];
		if (irs_offset)
		{
			a.x0 = a.x0 + g.ray_half_sep[0];
			a.y0 = a.y0 + g.ray_half_sep[1];
		}
		a.stepx = 2 * g.ray_half_sep[0];
		a.stepy = 2 * g.ray_half_sep[1];
		a.N1 = g.num_cells[0];
		a.N2 = g.num_cells[1];
		a.cyx = g.center_y[0];
		a.cyy = g.center_y[1];
		a.h1 = g.half_length_y[0];
		a.h2 = g.half_length_y[1];
		a.n1 = g.num_pixels[0];
		a.n2 = g.num_pixels[1];
		// the triangle queue packs a box row into 23 bits
		if (a.n2 >= (1 << 23)) throw std::runtime_error("num_pixels_y2 must be < 8388608");
		a.s1 = double(a.n1) / (2 * a.h1);
		a.s2 = double(a.n2) / (2 * a.h2);
		// ipm_functions.cuh:170
		a.factor = 2 * g.ray_half_sep[0] * g.ray_half_sep[1] * a.n1 * a.n2 / (2 * a.h1 * 2 * a.h2);
		return a;
	}
};

#define MLENS_TREE_ARGS(a)                \
	a.ncx = ncx_.ptr;                     \
	a.ncy = ncy_.ptr;                     \
	a.nhl = nhl_.ptr;                     \
	a.nchild0 = nchild0_.ptr;             \
	a.nbegin = nbegin_.ptr;               \
	a.ncount = ncount_.ptr;               \
	a.nnnbr = nnnbr_.ptr;                 \
	a.nnbr = nnbr_.ptr;                   \
	a.leaf_of_node = leaf_of_node_.ptr;   \
	a.local = local_.ptr;                 \
	a.p = p_;                             \
	a.sx = sx_.ptr;                       \
	a.sy = sy_.ptr;                       \
	a.sm = sm_.ptr;                       \
	a.leaf_geom = leaf_geom_.ptr;         \
	a.leaf_near_n = leaf_near_n_.ptr;     \
	a.near_d = leaf_near_d_.ptr;          \
	a.near_f = leaf_near_f_.ptr;          \
	a.poly_d = leaf_poly_d_.ptr;          \
	a.poly_f = leaf_poly_f_.ptr;

}  // namespace mlens
