// Shooting kernels: IPM cells (ipm_functions.cuh:112-230) and IRS rays
// (irs_functions.cuh:111-212), plus map utilities.
//
// Fast IPM kernel layout: a wave owns 64 corner columns (63 cells) and walks
// up a tile of rows; each lane deflects one corner per row and gets its right
// neighbour's corner by a lane shuffle, so every lattice corner is deflected
// once instead of four times (the reference deflects all four corners of each
// cell, ipm_functions.cuh:133-166). Lanes that sit in the same leaf evaluate
// it together: the leaf index is made wave-uniform with readfirstlane, so the
// leaf's star list and polynomial are read through scalar loads and used as
// scalar operands.
#include "shoot_common.hpp"
#include "wave_scan.hpp"

#include <algorithm>
#include <cstring>
#include <chrono>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace mlens
{

/* Deflect one point per lane. Lanes that share a leaf are evaluated together
   with the leaf index made wave-uniform. cache_* hold the lane's last leaf and
   a box slightly inside it; a point inside that box needs no new descent. */
struct LeafCache
{
	int node = -1;
	double x0 = 1, x1 = -1, y0 = 1, y1 = -1;
};

template <int PREC, bool MU>
__device__ __forceinline__ Deflection deflect(const ShootArgs& a, double zx, double zy, LeafCache& c)
{
	if (!(zx > c.x0 && zx < c.x1 && zy > c.y0 && zy < c.y1))
	{
		c.node = find_node(a, zx, zy);
		int L = a.leaf_of_node[c.node];
		if (L >= 0)
		{
			const double* g = a.leaf_geom + 4 * (long long)L;
			double hh = g[2] * (1 - 1e-12);
			c.x0 = g[0] - hh;
			c.x1 = g[0] + hh;
			c.y0 = g[1] - hh;
			c.y1 = g[1] + hh;
		}
		else
		{
			c.x0 = 1;
			c.x1 = -1;
			c.y0 = 1;
			c.y1 = -1;
		}
	}
	const int L = a.leaf_of_node[c.node];
	Deflection d{0, 0, 0};
	if (L >= 0)
	{
		bool pending = true;
		while (pending)
		{
			int u = __builtin_amdgcn_readfirstlane(L);
			// hide the equality L == u from the optimiser: otherwise it rewrites
			// u as the per-lane L inside the branch and the leaf tables are read
			// with per-lane vector loads instead of scalar loads
			asm volatile("" : "+s"(u));
			int same = (L == u);
			asm volatile("" : "+v"(same));
			if (same)
			{
				d = deflect_leaf<PREC, MU>(a, u, zx, zy);
				pending = false;
			}
		}
	}
	else
	{
		d = deflect_node<MU>(a, c.node, zx, zy);
	}
	return d;
}

/* ---------------------------------------------------------------------------
Triangle to pixels (polygon.cuh:349-467 semantics: the triangle's image-plane
area `factor` is spread over the pixels in proportion to the overlap, the
part outside the map is dropped, whole_area is taken before clipping).
--------------------------------------------------------------------------- */
/*
Per-lane accumulation of map contributions. Consecutive lattice rows of one
lane land in the same pixel most of the time, so contributions are summed in
two register slots (most recent first) and only an evicted slot becomes a
global atomic. On a multi-XCD MI355X device-scope atomics are carried out
beyond the per-XCD L2 caches, so every avoided atomic is a memory round trip.
The key is pixel index * 2 + map (0 = total or minima, 1 = saddles).
*/
/*
Map layouts. The fast IPM kernel accumulates into a map stored in 128-byte
tiles of 4 x 8 FP32 (4 x 4 FP64) pixels (width x height), tiles in row-major
order, and the map is folded into the row-major map afterwards. A small
triangle then spans one or two cache lines instead of one per pixel row, and
the atomics (carried out beyond the per-XCD L2 caches) cost per line touched.
On the 10000^2 map the FP32 kernel took 2.162 s with 4 x 8 tiles, 2.260 s
with 8 x 4, 2.475 s with 2 x 16 and 2.809 s with 16 x 2; the FP64 kernel
2.539 s with 4 x 4 and 2.666 s with 2 x 8.
*/
template <typename MapT>
struct MapTile
{
	static constexpr int W = 4, H = sizeof(MapT) == 4 ? 8 : 4;
};

// index of pixel (X, Y), X, Y >= 0, of a tiled map: the row's base plus the column's offset
template <typename MapT>
__device__ __forceinline__ long long tiled_row_base(int tiles_x, int Y)
{
	constexpr unsigned W = MapTile<MapT>::W, H = MapTile<MapT>::H;
	return (long long)(unsigned(Y) / H) * tiles_x * (W * H) + (unsigned(Y) % H) * W;
}

template <typename MapT>
__device__ __forceinline__ int tiled_col(int X)
{
	constexpr unsigned W = MapTile<MapT>::W, H = MapTile<MapT>::H;
	return int((unsigned(X) / W) * (W * H) + unsigned(X) % W);
}

template <typename MapT>
__device__ __forceinline__ long long tiled_index(int tiles_x, int X, int Y)
{
	return tiled_row_base<MapT>(tiles_x, Y) + tiled_col<MapT>(X);
}

/*
One row y of a map after the shot: map[y][x] += tiled[y][x] + s (= with ASSIGN), where s is the
running FP64 sum of span[y][0..x] (the span map has the tiled map's type and
layout; it may be null), and the tiled and span maps are cleared. A pixel that
no fully covered run reaches has s = 0 up to rounding, and every run weighs
more than 1e-4 (a triangle's weight per pixel is factor / area, and the kernel
drops triangles of area 1e4 * factor or more), while the residue is of order
2^-24 (FP32 spans) times the weights summed into the span cells before it along
the row; so |s| < 2^-16 is such a pixel, and it gets exactly nothing.
*/
template <typename MapT, bool ASSIGN>
__device__ __forceinline__ void fold_row(MapT* __restrict__ row, MapT* __restrict__ tiled, MapT* __restrict__ span, int n1,
	int tiles_x, int y)
{
	const int lane = threadIdx.x & 63;
	const long long rt = tiled_row_base<MapT>(tiles_x, y);
	double run = 0.0;
	for (int x0 = 0; x0 < n1; x0 += 64)
	{
		const int x = x0 + lane;
		const bool in = x < n1;
		const long long it = rt + tiled_col<MapT>(x);
		const MapT dm = (in && span) ? span[it] : MapT(0);
		const double d = double(dm);
		const MapT v = in ? tiled[it] : MapT(0);
		const double s = run + wave_prefix_sum(d);
		run = __shfl(s, 63);
		if (in)
		{
			const MapT add = v + (fabs(s) < 0x1p-16 ? MapT(0) : MapT(s));
			if (ASSIGN)
				row[x] = add;
			else if (add != MapT(0))
				row[x] += add;
			if (v != MapT(0)) tiled[it] = MapT(0);
			if (dm != MapT(0)) span[it] = MapT(0);
		}
	}
}

// the same for the map a kernel writes: tiled (TILED) or row-major
template <typename MapT, bool TILED>
__device__ __forceinline__ long long map_row_base(const ShootArgs& a, int Y)
{
	if constexpr (TILED)
		return tiled_row_base<MapT>(a.tiles_x, Y);
	else
		return (long long)a.n1 * Y;
}

template <typename MapT, bool TILED>
__device__ __forceinline__ int map_col(int X)
{
	if constexpr (TILED)
		return tiled_col<MapT>(X);
	else
		return X;
}

template <typename MapT, bool TILED>
__device__ __forceinline__ long long map_index(const ShootArgs& a, int X, int Y)
{
	return map_row_base<MapT, TILED>(a, Y) + map_col<MapT, TILED>(X);
}

template <typename MapT>
struct PixelAcc
{
	long long k0 = -1, k1 = -1;
	MapT v0 = 0, v1 = 0;
};

template <typename MapT, bool PARITY>
__device__ __forceinline__ void acc_flush_one(const ShootArgs& a, long long key, MapT v)
{
	if (key < 0) return;
	void* map = PARITY ? ((key & 1) ? a.msad : a.mmin) : a.map;
	if constexpr (std::is_same<MapT, int>::value)
		atomicAdd(reinterpret_cast<int*>(map) + (key >> 1), v);
	else
		unsafeAtomicAdd(reinterpret_cast<MapT*>(map) + (key >> 1), v);
}

template <typename MapT, bool PARITY>
__device__ __forceinline__ void acc_add(const ShootArgs& a, PixelAcc<MapT>& c, long long key, MapT v)
{
	if (key == c.k0)
	{
		c.v0 += v;
	}
	else if (key == c.k1)
	{
		MapT t = c.v1 + v;
		c.k1 = c.k0;
		c.v1 = c.v0;
		c.k0 = key;
		c.v0 = t;
	}
	else
	{
		acc_flush_one<MapT, PARITY>(a, c.k1, c.v1);
		c.k1 = c.k0;
		c.v1 = c.v0;
		c.k0 = key;
		c.v0 = v;
	}
}

template <typename MapT, bool PARITY>
__device__ __forceinline__ void acc_flush(const ShootArgs& a, PixelAcc<MapT>& c)
{
	acc_flush_one<MapT, PARITY>(a, c.k0, c.v0);
	acc_flush_one<MapT, PARITY>(a, c.k1, c.v1);
	c.k0 = c.k1 = -1;
	c.v0 = c.v1 = 0;
}

/*
General triangles (crossed by more than one pixel line) are rasterised by the
whole wave. A lane that meets one writes it, set up for per-pixel areas, into
its wave's queue in LDS; once the queue holds more than FLUSH triangles
the wave walks the pixel rows of all queued triangles, one row per lane, and
then the pixels of those rows, one per lane. A lane holding the 100-pixel
image of a demagnified region near a star then no longer stalls the other 63
(those images are a tenth of the general triangles and four fifths of their
pixels).
*/
// Entries per wave: 100 (12.7 KB with the ring and the mark array) for the
// single-map kernel, whose flushes then fill their passes better (10000^2 map:
// 1.695 s against 1.710 s with 79); 79 (10.7 KB) for the persistent batch
// kernel, whose 12 blocks per CU must all be resident at once (256 maps:
// 1.664 s with 79 entries, 1.866 s with 100)
constexpr int IPM_QCAP = 100;
constexpr int IPM_BATCH_QCAP = 79;

template <int CAP>
struct TriQueue
{
	static_assert(CAP <= 128, "a ring slot keeps the queue entry in 7 bits");
	static constexpr int FLUSH = CAP - 64;   // flush past this: an enqueue adds at most 64
	double x[3][CAP], y[3][CAP];             // ccw vertices relative to (X0, Y0)
	double k[3][CAP];                        // per edge: dx / dy (0 when dy = 0)
	double scale[CAP];                       // factor / triangle area
	int X0[CAP], Y0[CAP];
	int wb[CAP];                             // box width in pixels << 1 | map (0 = total or minima, 1 = saddles)
	int end[CAP];                            // box rows, then running total
	int ring_ey[4 * 64];                         // edge-crossed pixels awaiting their exact area: entry | map bit << 7 | row << 8
	int ring_x[4 * 64];                          //   and their map column
	int mark[4 * 64];                            // first slot of each entry / row starting in a pass; merge table
};

// a general triangle handed from the per-lane path to the wave queue
struct GenTri
{
	double x[3], y[3];   // vertices relative to (X0, Y0), in cell order
	int X0, Y0, w, h, bit;
};

/*
One triangle of a cell (polygon.cuh:349-467 semantics: its image-plane area
`factor` is spread over the pixels in proportion to the overlap; the part
outside the map is dropped; the whole area is taken before clipping).
Fast cases: inside one pixel (everything to that pixel), or cut by a single
pixel line: the vertex alone on one side of the line x = X cuts off a triangle
with area fraction t_b t_c, where t_b, t_c are the fractions of its two
edges on its own side (similar triangles); no area is ever differenced, so
slivers stay exact. Returns true when the triangle is left in g for the
wave's general rasteriser.
*/
template <int PREC, bool PARITY, typename MapT, bool TILED>
__device__ __forceinline__ bool ipm_triangle(const ShootArgs& a, PixelAcc<MapT>& acc, double p0x, double p0y, double p1x,
	double p1y, double p2x, double p2y, GenTri& g)
{
	// polygon.cuh:116-144 (difference form) and ipm_functions.cuh:176. A corner that is not finite
	// makes A NaN or infinite, so this test also drops every triangle with such a corner.
	double A = ((p1y - p0y) * p0x - (p1x - p0x) * p0y + (p2y - p1y) * p1x - (p2x - p1x) * p1y + (p0y - p2y) * p2x - (p0x - p2x) * p2y) / 2;
	if (!(fabs(A) < 10000 * a.factor) || A == 0) return false;
	const double minx = fmin(p0x, fmin(p1x, p2x)), maxx = fmax(p0x, fmax(p1x, p2x));
	const double miny = fmin(p0y, fmin(p1y, p2y)), maxy = fmax(p0y, fmax(p1y, p2y));
	if (minx >= a.n1 || maxx <= 0 || miny >= a.n2 || maxy <= 0) return false;
	const int map_bit = PARITY ? (A < 0 ? 0 : 1) : 0;
	const int X0 = int(floor(fmax(minx, 0.0)));
	const int Y0 = int(floor(fmax(miny, 0.0)));
	const int X1 = min(a.n1 - 1, int(floor(fmin(maxx, double(a.n1)))));
	const int Y1 = min(a.n2 - 1, int(floor(fmin(maxy, double(a.n2)))));
	const bool inside = minx >= 0 && maxx <= a.n1 && miny >= 0 && maxy <= a.n2;
	const MapT factor = MapT(a.factor);
	if (inside && X0 == X1 && Y0 == Y1)
	{
		acc_add<MapT, PARITY>(a, acc, map_index<MapT, TILED>(a, X0, Y0) * 2 + map_bit, factor);
		return false;
	}
	const bool xcut = inside && X1 == X0 + 1 && Y0 == Y1;
	const bool ycut = inside && Y1 == Y0 + 1 && X0 == X1;
	if (xcut || ycut)
	{
		// coordinate across the cut line, measured from it
		const double L = xcut ? double(X1) : double(Y1);
		double u0 = (xcut ? p0x : p0y) - L, u1 = (xcut ? p1x : p1y) - L, u2 = (xcut ? p2x : p2y) - L;
		const bool s0 = u0 < 0, s1 = u1 < 0, s2 = u2 < 0;   // below/left of the line
		const int nlow = int(s0) + int(s1) + int(s2);
		// the lone vertex: the one whose side has a single vertex
		const bool lone_low = (nlow == 1);
		double ua, ub, uc;
		if (s0 == lone_low)
		{
			ua = u0;
			ub = u1;
			uc = u2;
		}
		else if (s1 == lone_low)
		{
			ua = u1;
			ub = u2;
			uc = u0;
		}
		else
		{
			ua = u2;
			ub = u0;
			uc = u1;
		}
		// fractions t_b = ua / (ua - ub), t_c = ua / (ua - uc) of the edges from
		// the lone vertex to the line; their product in one division
		const MapT lone = MapT(double(a.factor) * (ua * ua) * dev::rcp64((ua - ub) * (ua - uc)));
		long long klow = map_index<MapT, TILED>(a, X0, Y0);
		long long khigh = xcut ? map_index<MapT, TILED>(a, X0 + 1, Y0) : map_index<MapT, TILED>(a, X0, Y0 + 1);
		long long klone = lone_low ? klow : khigh, kother = lone_low ? khigh : klow;
		acc_add<MapT, PARITY>(a, acc, klone * 2 + map_bit, lone);
		acc_add<MapT, PARITY>(a, acc, kother * 2 + map_bit, factor - lone);
		return false;
	}
	// general: FP64 vertices relative to the box origin; the wave's queue
	// orients them and sets up the edges
	const int w = X1 - X0 + 1, h = Y1 - Y0 + 1;
	g.x[0] = p0x - X0;
	g.y[0] = p0y - Y0;
	g.x[1] = p1x - X0;
	g.y[1] = p1y - Y0;
	g.x[2] = p2x - X0;
	g.y[2] = p2y - Y0;
	const double area2 = (g.x[1] - g.x[0]) * (g.y[2] - g.y[0]) - (g.x[2] - g.x[0]) * (g.y[1] - g.y[0]);
	if (area2 == 0) return false;
	g.X0 = X0;
	g.Y0 = Y0;
	g.w = w;
	g.h = h;
	g.bit = map_bit;
	return true;
}

// orders this wave's LDS accesses: the compiler keeps them in program order
// and the wave waits for its outstanding LDS operations
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

/*
Columns of row yy of queued triangle e: [xl, xr] are the pixels that lie
outside no edge's half-plane by more than 1e-9 of the edge's length, [il, ir]
those strictly inside all three (Green's theorem gives exactly 1 there).
With E(p) = dx (p_y - y_i) - dy (p_x - x_i), positive inside a
counter-clockwise triangle, a pixel is outside edge i when E's largest value
over its corners is below -tol and inside when its smallest is above tol.
Over a row, dx p_y + dy x_i - dx y_i ranges over [A, B]; the remaining
term -dy p_x bounds the column from the left (dy < 0) or the right (dy > 0),
and a horizontal edge admits the whole row or none of it.
*/
template <class Q>
__device__ __forceinline__ void tri_queue_row(const Q& q, int e, int yy, int& xl, int& xr, int& il, int& ir)
{
	const double wmax = double((q.wb[e] >> 1) - 1);
	const double Y = double(yy);
	double lo = 0.0, hi = wmax, ilo = 0.0, ihi = wmax;
#pragma unroll
	for (int i = 0; i < 3; i++)
	{
		const int n = (i + 1) % 3;
		const double xi = q.x[i][e], yi = q.y[i][e];
		const double dx = q.x[n][e] - xi, dy = q.y[n][e] - yi, idy = dy != 0 ? dev::rcp64(dy) : 0.0;
		const double tol = 1e-9 * (fabs(dx) + fabs(dy));
		const double base = dy * xi - dx * yi;
		const double A = fma(dx, dx > 0 ? Y : Y + 1.0, base);
		const double B = fma(dx, dx > 0 ? Y + 1.0 : Y, base);
		const double bA = (A - tol) * idy, bB = (B + tol) * idy;
		if (dy < 0)
		{
			// inside: A - dy X > tol, X > bA; not outside: B - dy (X + 1) >= -tol, X >= bB - 1
			ilo = fmax(ilo, floor(bA) + 1.0);
			lo = fmax(lo, ceil(bB) - 1.0);
		}
		else if (dy > 0)
		{
			// inside: A - dy (X + 1) > tol, X < bA - 1; not outside: B - dy X >= -tol, X <= bB
			ihi = fmin(ihi, ceil(bA) - 2.0);
			hi = fmin(hi, floor(bB));
		}
		else
		{
			if (!(A > tol)) ihi = -1.0;
			if (B < -tol) hi = -1.0;
		}
	}
	// lo, ilo >= 0 and hi, ihi <= w - 1 already; bound the other sides before converting
	xl = int(fmin(lo, wmax + 1.0));
	xr = int(fmax(hi, -1.0));
	il = int(fmin(ilo, wmax + 1.0));
	ir = int(fmax(ihi, -1.0));
}

// adds v to pixel pix of the map (PARITY: bit 0 = minima, 1 = saddles)
template <typename MapT, bool PARITY>
__device__ __forceinline__ void map_add(const ShootArgs& a, long long pix, int bit, MapT v)
{
	MapT* map = reinterpret_cast<MapT*>(PARITY ? (bit ? a.msad : a.mmin) : a.map);
	unsafeAtomicAdd(map + pix, v);
}

/*
A fully covered run [x0, x1] of pixel row Y gets weight w in every pixel. It is
kept as two differences in a span map of the map's type and tiles, +w at x0 and
-w at x1 + 1 (none past the map's last column); the fold adds each row's running
sum, taken in FP64, to the map. A run costs two atomics instead of one per
pixel, and a tile's line holds the run ends of 8 (FP32) or 4 (FP64) rows.
*/
template <typename MapT, bool PARITY>
__device__ __forceinline__ void span_add(const ShootArgs& a, int Y, int x0, int x1, int bit, double w)
{
	MapT* span = static_cast<MapT*>((PARITY && bit) ? a.span_sad : a.span);
	const long long rb = tiled_row_base<MapT>(a.tiles_x, Y);
	const MapT v = MapT(w);
	unsafeAtomicAdd(span + rb + tiled_col<MapT>(x0), v);
	if (x1 + 1 < a.n1) unsafeAtomicAdd(span + rb + tiled_col<MapT>(x1 + 1), -v);
}

/*
Exact areas (FP64, Green's theorem over the pixel's row) of ring slots
[start, start + count), one per lane. One pass holds 64 crossed pixels of a few
neighbouring triangles, and a sixth of them (10000^2 map) land on a pixel that
another lane of the pass also hits: small triangles around one pixel corner,
and the pixels along edges two triangles share. Atomics on one address
serialise, so those lanes are summed in LDS first: each lane writes its lane
number into the slot of its pixel (the flush's mark array is free here),
lanes whose slot holder has the same pixel add into the slot's cell, and only
the holders and the lanes that lost their slot to another pixel issue atomics.
128 slots of FP32 cells, or 64 of FP64, fill the mark array's 1 KB. On the
10000^2 map the merge took the mixed kernel from 2.062 to 1.744 s with 64
slots, and 128 slots took 2.5% more off (256 gained nothing); a second round
for the lanes that lost their slot made it slower (1.92 s).
*/
template <typename MapT, bool PARITY, bool TILED, class Q>
__device__ __forceinline__ void tri_queue_exact(const ShootArgs& a, Q& q, int start, int count)
{
	const int lane = threadIdx.x & 63;
	int slot = start + lane;
	asm volatile("" : "+v"(slot));
	// every lane takes part in the merge; lanes past count read stale ring
	// slots (entries and rows in range of the queue's arrays) and add nothing
	const bool valid = lane < count;
	const int ey = q.ring_ey[slot & 255], X = q.ring_x[slot & 255];
	const int e = ey & 127, yy = ey >> 8;
	double x[3], y[3], k[3];
#pragma unroll
	for (int i = 0; i < 3; i++)
	{
		x[i] = q.x[i][e];
		y[i] = q.y[i][e];
		k[i] = q.k[i][e];
	}
	const double ar = dev::tri_pixel_area(x, y, k, double(X - q.X0[e]), double(yy));
	const bool live = valid && ar > 0;
	const int bit = (ey >> 7) & 1;
	const long long pix = live ? map_index<MapT, TILED>(a, X, q.Y0[e] + yy) : 0;
	MapT v = MapT(q.scale[e] * ar);
	constexpr int SLOTS = sizeof(MapT) == 4 ? 128 : 64;
	static_assert(SLOTS * (1 + int(sizeof(MapT)) / 4) <= 4 * 64, "hash slots and cells exceed the mark array");
	int* holder = q.mark;
	MapT* cell = reinterpret_cast<MapT*>(q.mark + SLOTS);
	const long long key = pix * 2 + bit;   // compared in full: maps may exceed 2^31 pixels
	// slot = the low bits of the tiled pixel index: distinct pixels of two neighbouring tiles never share one
	const int h = int((PARITY ? key : pix) & (SLOTS - 1));
	if (live) holder[h] = lane;
	wave_lds_sync();
	const int w = live ? holder[h] : lane;
	const long long kw = __shfl(key, w);   // outside any condition: a shuffle reads nothing from inactive lanes
	const bool merged = live && w != lane && kw == key;
	if (live && w == lane) cell[h] = v;
	wave_lds_sync();
	if (merged) atomicAdd(&cell[h], v);
	wave_lds_sync();
	if (live && w == lane) v = cell[h];
	if (live && !merged) map_add<MapT, PARITY>(a, pix, bit, v);
}

/*
Rasterise every queued triangle. Rows of the concatenated boxes go 64 at a
time to the lanes, which find their row's candidate and fully covered
columns and hand the fully covered run to the span map (span_add); the
remaining candidate pixels of those 64 rows, the ones an edge may cross, then
go to the lanes, 64 at a time, and from there to the ring, where they get
their exact areas in full passes of 64. A lane finds its entry (or row) as the
last one that starts at or before its own index: every entry (row) starting in
the current pass marks its first slot in LDS, and a max-scan over the slots
carries each mark forward, starting from the entry (row) the previous pass
ended in. rh / rt: next ring slot to process / to fill.
*/
template <typename MapT, bool PARITY, bool TILED, class Q>
__device__ __forceinline__ void tri_queue_flush(const ShootArgs& a, Q& q, int qn)
{
	const double factor = a.factor;
	const int lane = threadIdx.x & 63;
	wave_lds_sync();
	// set-up: counter-clockwise order, edge slopes, factor / area
	for (int base = 0; base < qn; base += 64)
	{
		int e = base + lane;
		asm volatile("" : "+v"(e));
		if (e < qn)
		{
			double x[3], y[3];
#pragma unroll
			for (int i = 0; i < 3; i++)
			{
				x[i] = q.x[i][e];
				y[i] = q.y[i][e];
			}
			double area2 = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
			if (area2 < 0)
			{
				const double sx = x[1], sy = y[1];
				x[1] = x[2];
				y[1] = y[2];
				x[2] = sx;
				y[2] = sy;
				area2 = -area2;
				q.x[1][e] = x[1];
				q.y[1][e] = y[1];
				q.x[2][e] = x[2];
				q.y[2][e] = y[2];
			}
#pragma unroll
			for (int i = 0; i < 3; i++)
			{
				const int n = (i + 1) % 3;
				const double dy = y[n] - y[i];
				const double idy = dy != 0 ? dev::rcp64(dy) : 0.0;
				q.k[i][e] = (x[n] - x[i]) * idy;
			}
			q.scale[e] = factor * dev::rcp64(area2 * 0.5);
		}
	}
	int total = 0;
	for (int base = 0; base < qn; base += 64)
	{
		int e = base + lane;
		asm volatile("" : "+v"(e));
		const int c = wave_prefix_sum(e < qn ? q.end[e] : 0);
		if (e < qn) q.end[e] = total + c;
		total += __builtin_amdgcn_readlane(c, 63);
	}
	wave_lds_sync();
	// first rows of this lane's entries (lane, lane + 64), or past the end
	const int s0 = lane < qn ? (lane > 0 ? q.end[lane - 1] : 0) : total;
	const int s1 = lane + 64 < qn ? q.end[lane + 63] : total;
	int rh = 0, rt = 0;
	int ecarry = 0;
	for (int r0 = 0; r0 < total; r0 += 64)
	{
		// the entry of row r0 + lane
		q.mark[lane] = -1;
		if (s0 >= r0 && s0 < r0 + 64) q.mark[s0 - r0] = lane;
		if (s1 >= r0 && s1 < r0 + 64) q.mark[s1 - r0] = lane + 64;
		wave_lds_sync();
		int e = max(wave_max_scan(q.mark[lane]), ecarry);
		ecarry = __builtin_amdgcn_readlane(e, 63);
		e = min(e, qn - 1);
		// this lane's row: its candidate columns, cnt of them, are the cl columns
		// from xa (absolute) and the rest after a gap (the fully covered run)
		const int r = r0 + lane;
		int xa = 0, cnt = 0, cl = 0, gap = 0, ey = 0;
		if (r < total)
		{
			const int start = e > 0 ? q.end[e - 1] : 0;
			const int yy = r - start;
			if (yy >= 0 && yy < q.end[e] - start)
			{
				int xl, xr, il, ir;
				tri_queue_row(q, e, yy, xl, xr, il, ir);
				const int X0 = q.X0[e], bit = q.wb[e] & 1;
				cnt = max(xr - xl + 1, 0);
				cl = cnt;
				il = max(il, xl);
				ir = min(ir, xr);
				if (il <= ir)
				{
					span_add<MapT, PARITY>(a, q.Y0[e] + yy, X0 + il, X0 + ir, bit, q.scale[e]);
					cl = il - xl;
					gap = ir - il + 1;
					cnt -= gap;
				}
				xa = X0 + xl;
				ey = e | (bit << 7) | (yy << 8);
			}
		}
		int off = wave_prefix_sum(cnt);
		const int tot = __builtin_amdgcn_readlane(off, 63);
		const int lim = off;   // candidate k of this row: off - cnt <= k < lim
		off -= cnt;
		const int xbase = xa - off;   // column of candidate k: xbase + k, plus gap from k = thr on
		const int thr = off + cl;
		int jcarry = 0;
		// candidate pixel k of row j: the row's data, fetched from the row's lane
		struct Cand
		{
			int k, x, lim, ey;
		};
		auto fetch = [&](int k, int j) {
			// every lane takes part in every shuffle: a lane shuffle reads nothing
			// from a lane that is inactive, so none may sit in a conditional arm
			const int xj = __shfl(xbase, j), tj = __shfl(thr, j), gj = __shfl(gap, j);
			Cand c;
			c.k = k;
			c.x = xj + k + (k >= tj ? gj : 0);
			c.lim = __shfl(lim, j);
			c.ey = __shfl(ey, j);
			return c;
		};
		// every candidate joins the ring
		auto pixel = [&](const Cand& c) {
			const bool crossed = c.k < tot && c.k < c.lim;
			const unsigned long long m = __ballot(crossed);
			if (crossed)
			{
				int pos = rt + int(__builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u)));
				asm volatile("" : "+v"(pos));
				q.ring_ey[pos & 255] = c.ey;
				q.ring_x[pos & 255] = c.x;
			}
			rt += __popcll(m);
		};
		// two passes of 64 candidates at a time: their row look-ups are independent
		// until the carry, so their LDS, DPP and shuffle latencies overlap. The ring
		// holds fewer than 64 slots before and at most 128 more after.
		for (int k0 = 0; k0 < tot; k0 += 128)
		{
			q.mark[lane] = -1;
			q.mark[lane + 64] = -1;
			if (cnt > 0 && off >= k0 && off < k0 + 128) q.mark[off - k0] = lane;
			wave_lds_sync();
			const int ma = q.mark[lane], mb = q.mark[lane + 64];
			const int ja = max(wave_max_scan(ma), jcarry);
			const int jb = max(wave_max_scan(mb), __builtin_amdgcn_readlane(ja, 63));
			jcarry = __builtin_amdgcn_readlane(jb, 63);
			const Cand ca = fetch(k0 + lane, ja), cb = fetch(k0 + 64 + lane, jb);
			pixel(ca);
			if (k0 + 64 < tot) pixel(cb);
			while (rt - rh >= 64)
			{
				wave_lds_sync();
				tri_queue_exact<MapT, PARITY, TILED>(a, q, rh, 64);
				rh += 64;
			}
		}
	}
	if (rt > rh)
	{
		wave_lds_sync();
		tri_queue_exact<MapT, PARITY, TILED>(a, q, rh, rt - rh);
	}
	wave_lds_sync();
}

// wave-wide: lanes with gen set append g to the queue
template <class Q>
__device__ __forceinline__ void tri_queue_push(Q& q, int& qn, bool gen, const GenTri& g)
{
	const unsigned long long m = __ballot(gen);
	if (m == 0) return;
	if (gen)
	{
		int pos = qn + int(__builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u)));
		// keep the index opaque: otherwise the lane part of every field's
		// address is hoisted out of the shooting loop, one register each
		asm volatile("" : "+v"(pos));
#pragma unroll
		for (int i = 0; i < 3; i++)
		{
			q.x[i][pos] = g.x[i];
			q.y[i][pos] = g.y[i];
		}
		q.X0[pos] = g.X0;
		q.Y0[pos] = g.Y0;
		q.wb[pos] = (g.w << 1) | g.bit;
		q.end[pos] = g.h;
	}
	qn += __popcll(m);
}

// waves per block: one for the single-map kernel (2.39 s on the 10000^2 map,
// 2.409 s with two, 2.428 s with four) and for the persistent batch kernel, so
// that no wave waits for the slowest wave of its block
constexpr int IPM_WAVES = 1;
constexpr int IPM_TILE_W = 63 * IPM_WAVES;
constexpr int IPM_BATCH_WAVES = 1;
constexpr int IPM_BATCH_TILE_W = 63 * IPM_BATCH_WAVES;
constexpr int IPM_RESIDENT_WAVES_PER_CU = 12;   // 3 per SIMD: up to 168 VGPRs and a 12.7 KB wave queue
constexpr int IPM_TILE_H = 64;       // lattice rows a wave walks: the shortest tile
constexpr int IPM_TILE_H_MAX = 1024;  // and the tallest

/* Rows per tile for a launch of ntx tile columns over `rows` lattice rows: the
   tallest power of two in [IPM_TILE_H, IPM_TILE_H_MAX] that still gives at
   least eight blocks per resident block slot. A tall tile deflects
   its first row once per 1024 rows instead of once per 64 and flushes its
   queue less often (10000^2 map: 2.475 s at 64 rows, 2.428 s at 1024); a short
   one keeps a small lattice spread over the whole device. */
static inline int ipm_tile_rows(long long ntx, long long rows, int num_cus)
{
	const long long want = 8LL * (IPM_RESIDENT_WAVES_PER_CU / IPM_WAVES) * (num_cus > 0 ? num_cus : 256);
	int h = IPM_TILE_H_MAX;
	while (h > IPM_TILE_H && ntx * ((rows + h - 1) / h) < want) h >>= 1;
	return h;
}

// One tile: this wave's 63 cells of lattice columns tile_x * 63 WAVES + wave * 63 + [0, 63),
// cell rows [J0, J1). Every lane of the wave must call it.
template <int PREC, bool PARITY, bool TILED, int WAVES, class Q>
__device__ __forceinline__ void ipm_tile(const ShootArgs& a, Q& q, long long tile_x, long long J0, long long J1)
{
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const long long I = tile_x * (63 * WAVES) + wave * 63 + lane;
	const long long Ic = I < a.N1 ? I : a.N1;
	const double zx = lattice_at(a.x0, a.stepx, Ic);
	const bool cell_ok = lane < 63 && I < a.N1;
	LeafCache cache;
	using MapT = typename std::conditional<PREC == PREC_MIXED, float, double>::type;
	PixelAcc<MapT> acc;
	int qn = 0;

	auto corner = [&](long long J, double& px, double& py) {
		double zy = lattice_at(a.y0, a.stepy, J);
		Deflection d = deflect<PREC, false>(a, zx, zy, cache);
		// ipm_functions.cuh:159-165 and array_functions.cuh:70-75
		px = ((d.wx - a.cyx) + a.h1) * a.s1;
		py = double(a.n2) - ((d.wy - a.cyy) + a.h2) * a.s2;
	};

	double bx, by;
	corner(J0, bx, by);
	double brx = __shfl_down(bx, 1), bry = __shfl_down(by, 1);
	for (long long J = J0 + 1; J <= J1; J++)
	{
		double tx, ty;
		corner(J, tx, ty);
		double trx = __shfl_down(tx, 1), try_ = __shfl_down(ty, 1);
		// cell (I, J-1): y0 = upper right, y1 = upper left, y2 = lower left, y3 = lower right;
		// triangles (y0, y1, y2) and (y2, y3, y0). One call site each for the
		// triangle, the queue and its flush keeps the register footprint small.
#pragma unroll 1
		for (int t = 0; t < 2; t++)
		{
			const double ax = t ? bx : trx, ay = t ? by : try_;
			const double bxx = t ? brx : tx, byy = t ? bry : ty;
			const double cx = t ? trx : bx, cy = t ? try_ : by;
			GenTri g;
			const bool gen = cell_ok && ipm_triangle<PREC, PARITY, MapT, TILED>(a, acc, ax, ay, bxx, byy, cx, cy, g);
			tri_queue_push(q, qn, gen, g);
			if (qn > Q::FLUSH || (qn > 0 && t == 1 && J == J1))
			{
				tri_queue_flush<MapT, PARITY, TILED>(a, q, qn);
				qn = 0;
			}
		}
		bx = tx;
		by = ty;
		brx = trx;
		bry = try_;
	}
	acc_flush<MapT, PARITY>(a, acc);
}

// Three waves per SIMD: 136 VGPRs without spills, and 12.7 KB of LDS per wave for a
// 100-entry queue and a 128-slot merge table. Four waves (128 VGPRs, about 25
// spilled) ran as fast with a 64-slot table and a 79-entry queue, the most that
// fit in 10 KB, and gained nothing from a larger table (10000^2 map: 1.698 s at
// three waves against 1.745 s at four)
template <int PREC, bool PARITY>
__global__ __launch_bounds__(64 * IPM_WAVES, 3) void k_ipm_fast(ShootArgs a, long long row0, long long row1,
	MLENS_TABLE_PARAMS, void* __restrict__ map, void* __restrict__ mmin, void* __restrict__ msad,
	void* __restrict__ span, void* __restrict__ span_sad, int tile_h)
{
	MLENS_TABLE_REBIND(a)
	a.map = map;
	a.mmin = mmin;
	a.msad = msad;
	a.span = span;
	a.span_sad = span_sad;
	const long long J0 = row0 + (long long)blockIdx.y * tile_h;
	if (J0 >= row1) return;
	const long long J1 = min(J0 + tile_h, row1);
	__shared__ TriQueue<IPM_QCAP> queues[IPM_WAVES];
	ipm_tile<PREC, PARITY, true, IPM_WAVES>(a, queues[threadIdx.x >> 6], blockIdx.x, J0, J1);
}

/*
Lattice cells that can reach a window of the map, for checking a window of a
map too large to shoot on the host: every cell whose corners' pixel box (over
its finite corners, so a superset of each surviving triangle's box) meets the
window [wx0, wx1) x [wy0, wy1) in pixel coordinates (row 0 at the top). The
walk is the mixed-precision lattice walk of k_ipm_fast. One record per lattice
row and 63-cell tile that holds such a cell: (row, first column, mask of the
tile's cells). count ends as the number of records; those beyond cap are not
written.
*/
__global__ __launch_bounds__(64 * IPM_WAVES) void k_window_cells(ShootArgs a, long long row0, long long row1,
	MLENS_TABLE_PARAMS, double wx0, double wx1, double wy0, double wy1, long long* __restrict__ rec, long long cap,
	unsigned long long* __restrict__ count)
{
	MLENS_TABLE_REBIND(a)
	const long long J0 = row0 + (long long)blockIdx.y * IPM_TILE_H;
	if (J0 >= row1) return;
	const long long J1 = min(J0 + IPM_TILE_H, row1);
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const long long Ibase = (long long)blockIdx.x * IPM_TILE_W + wave * 63;
	const long long I = Ibase + lane;
	const long long Ic = I < a.N1 ? I : a.N1;
	const double zx = lattice_at(a.x0, a.stepx, Ic);
	const bool cell_ok = lane < 63 && I < a.N1;
	LeafCache cache;
	auto corner = [&](long long J, double& px, double& py) {
		double zy = lattice_at(a.y0, a.stepy, J);
		Deflection d = deflect<PREC_MIXED, false>(a, zx, zy, cache);
		px = ((d.wx - a.cyx) + a.h1) * a.s1;
		py = double(a.n2) - ((d.wy - a.cyy) + a.h2) * a.s2;
	};
	double bx, by;
	corner(J0, bx, by);
	double brx = __shfl_down(bx, 1), bry = __shfl_down(by, 1);
	for (long long J = J0 + 1; J <= J1; J++)
	{
		double tx, ty;
		corner(J, tx, ty);
		const double trx = __shfl_down(tx, 1), try_ = __shfl_down(ty, 1);
		// fmin / fmax skip NaN corners; all four NaN leaves NaN, which meets nothing
		const double minx = fmin(fmin(tx, trx), fmin(bx, brx)), maxx = fmax(fmax(tx, trx), fmax(bx, brx));
		const double miny = fmin(fmin(ty, try_), fmin(by, bry)), maxy = fmax(fmax(ty, try_), fmax(by, bry));
		const bool hit = cell_ok && minx < wx1 && maxx > wx0 && miny < wy1 && maxy > wy0;
		const unsigned long long m = __ballot(hit);
		if (lane == 0 && m)
		{
			const unsigned long long slot = atomicAdd(count, 1ull);
			if ((long long)slot < cap)
			{
				rec[3 * slot] = J - 1;
				rec[3 * slot + 1] = Ibase;
				rec[3 * slot + 2] = (long long)m;
			}
		}
		bx = tx;
		by = ty;
		brx = trx;
		bry = try_;
	}
}

long long Engine::window_cells(const LensParams& lens, const GridParams& g, int wx0, int wy0, int ww, int wh,
	double margin, long long* rec, long long cap)
{
	if (ww < 1 || wh < 1 || cap < 0 || !(margin >= 0))
		throw HipError("window_cells: the window needs a positive size and margin >= 0", __FILE__, __LINE__);
	ShootArgs a = ShootArgsBuilder::make(lens, g, false);
	MLENS_TREE_ARGS(a)
	DeviceBuffer<long long> drec;
	if (cap > 0) drec.alloc(size_t(3 * cap));
	DeviceBuffer<unsigned long long> dcount;
	dcount.alloc(1);
	MLENS_HIP_CHECK(hipMemsetAsync(dcount.ptr, 0, sizeof(unsigned long long), stream_));
	dim3 grid(unsigned((a.N1 + IPM_TILE_W - 1) / IPM_TILE_W), unsigned((a.N2 + IPM_TILE_H - 1) / IPM_TILE_H));
	k_window_cells<<<grid, 64 * IPM_WAVES, 0, stream_>>>(a, 0, a.N2, MLENS_TABLE_ARGS(a), wx0 - margin, wx0 + ww + margin,
		wy0 - margin, wy0 + wh + margin, drec.ptr, cap, dcount.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	unsigned long long n = 0;
	MLENS_HIP_CHECK(hipMemcpyAsync(&n, dcount.ptr, sizeof(n), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	const long long got = (long long)n < cap ? (long long)n : cap;
	if (got > 0) MLENS_HIP_CHECK(hipMemcpy(rec, drec.ptr, size_t(3 * got) * sizeof(long long), hipMemcpyDeviceToHost));
	return (long long)n;
}

/* ---------------------------------------------------------------------------
Many maps in one launch. A persistent grid takes (instance, tile) work items
from a global counter: item t covers tile t * stride of the concatenation of
every instance's tiles (stride > 1 samples the batch for the autotuner). The
instance of a tile comes from a binary search over the running tile counts;
its arguments, including its own tables and map, are read from device memory.
--------------------------------------------------------------------------- */
struct BatchMap
{
	float* map;         // row-major n2 x n1, written by the fold
	float* tiled;       // tiled accumulation map the shooting kernel writes (MapTile<float>)
	float* span;        // its span map (span_add), laid out as the tiled map
	long long npix;     // n1 * n2
	long long ntiled;   // elements of the tiled map and of the span map
	int n1, n2, tiles_x;
};

// Three waves per SIMD (141 VGPRs, no spills): the batch's maps of 0.25-16 MB
// stay in the caches, and at four waves (128 VGPRs) the spills cost more than
// the fourth wave hides (256 maps: 1.669 s at 12 blocks per CU, 1.890 s at 16).
// Blocks take tiles from a global counter and never wait for each other.
__global__ __launch_bounds__(64 * IPM_BATCH_WAVES, 3) void k_ipm_batch(const ShootArgs* __restrict__ args,
	const long long* __restrict__ tile_start, int ninst, int tile_h, long long total, long long stride,
	unsigned long long* __restrict__ counter)
{
	__shared__ TriQueue<IPM_BATCH_QCAP> queues[IPM_BATCH_WAVES];
	__shared__ long long next;
	for (;;)
	{
		__syncthreads();   // every thread has read the previous item
		if (threadIdx.x == 0) next = (long long)atomicAdd(counter, 1ull) * stride;
		__syncthreads();
		const long long t = next;
		if (t >= total) break;
		int lo = 0, hi = ninst - 1;
		while (lo < hi)
		{
			const int mid = (lo + hi + 1) >> 1;
			if (tile_start[mid] <= t)
				lo = mid;
			else
				hi = mid - 1;
		}
		const ShootArgs a = args[lo];
		const long long local = t - tile_start[lo];
		const long long ntx = (a.N1 + IPM_BATCH_TILE_W - 1) / IPM_BATCH_TILE_W;
		const long long ty = local / ntx, tx = local - ty * ntx;
		const long long J0 = ty * tile_h, J1 = min(J0 + (long long)tile_h, a.N2);
		ipm_tile<PREC_MIXED, false, true, IPM_BATCH_WAVES>(a, queues[threadIdx.x >> 6], tx, J0, J1);
	}
}

// zero every instance's tiled and span maps (blockIdx.y = instance), its histogram and the work counter
__global__ void k_batch_zero(const BatchMap* __restrict__ maps, unsigned* hist, unsigned long long* counter)
{
	const int inst = blockIdx.y;
	if (blockIdx.x == 0)
	{
		for (int k = threadIdx.x; k < BATCH_HIST_BINS; k += blockDim.x) hist[(long long)inst * BATCH_HIST_BINS + k] = 0;
		if (inst == 0 && threadIdx.x == 0) *counter = 0;
	}
	const BatchMap m = maps[inst];
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m.ntiled; i += (long long)gridDim.x * blockDim.x)
	{
		m.tiled[i] = 0.0f;
		m.span[i] = 0.0f;
	}
}

// every instance's row-major map from its tiled and span maps (blockIdx.y = instance, one wave per
// row, blocks of four rows striding over the instance's rows); both are cleared
__global__ __launch_bounds__(256) void k_batch_fold(const BatchMap* __restrict__ maps)
{
	const BatchMap m = maps[blockIdx.y];
	for (int y = blockIdx.x * 4 + (threadIdx.x >> 6); y < m.n2; y += gridDim.x * 4)
		fold_row<float, true>(m.map + (long long)m.n1 * y, m.tiled, m.span, m.n1, m.tiles_x, y);
}

// 128-bin histogram of log10(magnification) over [-2, 2] per instance; values
// below 10^-2 (zero included) count in the first bin, above 10^2 in the last
__global__ void k_batch_hist(const BatchMap* __restrict__ maps, unsigned* hist)
{
	__shared__ unsigned h[BATCH_HIST_BINS];
	for (int k = threadIdx.x; k < BATCH_HIST_BINS; k += blockDim.x) h[k] = 0;
	__syncthreads();
	const BatchMap m = maps[blockIdx.y];
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m.npix; i += (long long)gridDim.x * blockDim.x)
	{
		const float v = m.map[i];
		int b = v > 0.0f ? int(floorf((log10f(v) + 2.0f) * (BATCH_HIST_BINS / 4.0f))) : 0;
		b = min(max(b, 0), BATCH_HIST_BINS - 1);
		atomicAdd(&h[b], 1u);
	}
	__syncthreads();
	for (int k = threadIdx.x; k < BATCH_HIST_BINS; k += blockDim.x)
		if (h[k]) atomicAdd(&hist[(long long)blockIdx.y * BATCH_HIST_BINS + k], h[k]);
}

/* ---------------------------------------------------------------------------
Reference-faithful FP64 IPM kernel: one thread per cell, four corners each
with its own tree descent, Smith divisions in alpha_star, the reference's
polygon clipping with log/exp edge intersections (polygon.cuh:150-467).
It is the plain-FP64 baseline the fast kernels are measured against.
--------------------------------------------------------------------------- */
namespace refpoly
{
constexpr int MAXS = 7;
struct Poly
{
	double x[MAXS], y[MAXS];
	int n;
};
__device__ inline void add(Poly& P, double x, double y)
{
	if (P.n >= MAXS)
	{
		P.n = 0;
		return;
	}
	P.x[P.n] = x;
	P.y[P.n] = y;
	P.n++;
}
__device__ inline double area(const Poly& P)
{
	if (P.n < 3) return 0;
	double r = 0;
	for (int i = 0; i < P.n; i++)
	{
		int j = (i + 1) % P.n;
		r += (P.y[j] - P.y[i]) * P.x[i] - (P.x[j] - P.x[i]) * P.y[i];
	}
	return r / 2;
}
__device__ inline double sgn(double v) { return v < -0.0 ? -1 : (v > 0 ? 1 : 0); }
__device__ inline double yint(double X, double ax, double ay, double bx, double by)
{
	double dy = by - ay;
	if (dy == 0) return ay;
	double ldy = log(fabs(dy)), dx = bx - ax, ldx = log(fabs(dx));
	double lt = log(fabs(X - ax)) - ldx;
	return ay + sgn(X - ax) * sgn(dx) * sgn(dy) * exp(lt + ldy);
}
__device__ inline double xint(double Y, double ax, double ay, double bx, double by)
{
	double dx = bx - ax;
	if (dx == 0) return ax;
	double ldx = log(fabs(dx)), dy = by - ay, ldy = log(fabs(dy));
	double lt = log(fabs(Y - ay)) - ldy;
	return ax + sgn(Y - ay) * sgn(dy) * sgn(dx) * exp(lt + ldx);
}
// keep = the side kept in P; returns the other side
__device__ inline void clip_x(Poly& P, double X, bool clip_left, Poly& out)
{
	Poly L, R;
	L.n = R.n = 0;
	for (int i = 0; i < P.n; i++)
	{
		int j = (i + 1) % P.n;
		double ax = P.x[i], ay = P.y[i], bx = P.x[j], by = P.y[j];
		if (ax < X)
		{
			add(L, ax, ay);
			if (bx > X)
			{
				double y = yint(X, ax, ay, bx, by);
				add(L, X, y);
				add(R, X, y);
			}
		}
		else if (ax > X)
		{
			add(R, ax, ay);
			if (bx < X)
			{
				double y = yint(X, ax, ay, bx, by);
				add(L, X, y);
				add(R, X, y);
			}
		}
		else
		{
			add(L, ax, ay);
			add(R, ax, ay);
		}
	}
	if (clip_left)
	{
		P = R;
		out = L;
	}
	else
	{
		P = L;
		out = R;
	}
}
__device__ inline void clip_y(Poly& P, double Y, bool clip_bottom, Poly& out)
{
	Poly B, T;
	B.n = T.n = 0;
	for (int i = 0; i < P.n; i++)
	{
		int j = (i + 1) % P.n;
		double ax = P.x[i], ay = P.y[i], bx = P.x[j], by = P.y[j];
		if (ay < Y)
		{
			add(B, ax, ay);
			if (by > Y)
			{
				double x = xint(Y, ax, ay, bx, by);
				add(B, x, Y);
				add(T, x, Y);
			}
		}
		else if (ay > Y)
		{
			add(T, ax, ay);
			if (by < Y)
			{
				double x = xint(Y, ax, ay, bx, by);
				add(B, x, Y);
				add(T, x, Y);
			}
		}
		else
		{
			add(B, ax, ay);
			add(T, ax, ay);
		}
	}
	if (clip_bottom)
	{
		P = T;
		out = B;
	}
	else
	{
		P = B;
		out = T;
	}
}
__device__ inline double minx(const Poly& P) { double m = P.x[0]; for (int i = 1; i < P.n; i++) m = fmin(m, P.x[i]); return m; }
__device__ inline double maxx(const Poly& P) { double m = P.x[0]; for (int i = 1; i < P.n; i++) m = fmax(m, P.x[i]); return m; }
__device__ inline double miny(const Poly& P) { double m = P.y[0]; for (int i = 1; i < P.n; i++) m = fmin(m, P.y[i]); return m; }
__device__ inline double maxy(const Poly& P) { double m = P.y[0]; for (int i = 1; i < P.n; i++) m = fmax(m, P.y[i]); return m; }

// polygon.cuh:349-467
__device__ void allocate(Poly P, double factor, double* pixels, int n1, int n2)
{
	if (P.n < 3) return;
	if (minx(P) >= n1 || maxx(P) <= 0 || miny(P) >= n2 || maxy(P) <= 0) return;
	double whole = fabs(area(P));
	if (whole == 0) return;
	Poly tmp;
	if (minx(P) < 0) clip_x(P, 0, true, tmp);
	if (maxx(P) > n1) clip_x(P, n1, false, tmp);
	if (miny(P) < 0) clip_y(P, 0, true, tmp);
	if (maxy(P) > n2) clip_y(P, n2, false, tmp);
	if (P.n < 3) return;
	int xmin = int(minx(P)), xmax = int(maxx(P));
	if (xmax == n1) xmax--;
	for (int x = xmin; x < xmax; x++)
	{
		Poly top;
		clip_x(P, x + 1, true, top);
		if (top.n < 3) continue;
		int ymin = int(miny(top)), ymax = int(maxy(top));
		if (ymax == n2) ymax--;
		for (int y = ymin; y < ymax; y++)
		{
			Poly bot;
			clip_y(top, y + 1, true, bot);
			if (bot.n < 3) continue;
			unsafeAtomicAdd(&pixels[(long long)n1 * y + x], factor * fabs(area(bot)) / whole);
		}
		unsafeAtomicAdd(&pixels[(long long)n1 * ymax + x], factor * fabs(area(top)) / whole);
	}
	if (P.n < 3) return;
	int ymin = int(miny(P)), ymax = int(maxy(P));
	if (ymax == n2) ymax--;
	for (int y = ymin; y < ymax; y++)
	{
		Poly bot;
		clip_y(P, y + 1, true, bot);
		if (bot.n < 3) continue;
		unsafeAtomicAdd(&pixels[(long long)n1 * y + xmax], factor * fabs(area(bot)) / whole);
	}
	unsafeAtomicAdd(&pixels[(long long)n1 * ymax + xmax], factor * fabs(area(P)) / whole);
}
}  // namespace refpoly

// ipm_functions.cuh:34-46 with alpha_star.cuh:18-41 (Smith division) and alpha_local.cuh:18-36
__device__ Deflection deflect_reference(const ShootArgs& a, double zx, double zy)
{
	int node = find_node(a, zx, zy);
	dcx as = {0, 0};
	auto range = [&](int b, int n) {
		for (int s = b; s < b + n; s++) as = dev::add(as, dev::cdiv(dcx{a.sm[s], 0}, dcx{zx - a.sx[s], zy - a.sy[s]}));
	};
	for (int j = 0; j < a.nnnbr[node]; j++)
	{
		int nb = a.nnbr[8 * node + j];
		range(a.nbegin[nb], a.ncount[nb]);
	}
	range(a.nbegin[node], a.ncount[node]);
	as = dev::conj(dev::scale(as, a.theta2));
	const double h = a.nhl[node];
	dcx dz = dcx{(zx - a.ncx[node]) / h, (zy - a.ncy[node]) / h};
	dcx al = {0, 0};
	const double* c = a.local + 2 * (long long)node * (a.p + 1);
	for (int i = a.p; i >= 1; i--)
	{
		al = dev::mul(al, dz);
		al = dev::add(al, dcx{c[2 * i] * i, c[2 * i + 1] * i});
	}
	al = dev::scale(al, a.theta2);
	al = dev::scale(al, 1.0 / h);
	al = dev::conj(al);
	dcx sm = dev::alpha_smooth(dcx{zx, zy}, a.kappastar, a.rectangular, a.approx, a.taylor_smooth, dcx{a.corner_re, a.corner_im});
	Deflection d;
	d.wx = (1 - a.kappa) * zx + a.gamma * zx - as.re - al.re - sm.re;
	d.wy = (1 - a.kappa) * zy - a.gamma * zy - as.im - al.im - sm.im;
	d.mu = 0;
	return d;
}

template <bool PARITY>
__global__ __launch_bounds__(256) void k_ipm_reference(ShootArgs a, long long row0, long long row1)
{
	long long i = (long long)blockIdx.x * 16 + threadIdx.x;
	long long j = row0 + (long long)blockIdx.y * 16 + threadIdx.y;
	if (i >= a.N1 || j >= row1) return;
	// ipm_functions.cuh:125-130 (a.x0, a.y0 hold center_x - half_length_x)
	double hx = a.stepx / 2, hy = a.stepy / 2;
	double zx = a.x0 + hx + a.stepx * double(i);
	double zy = a.y0 + hy + a.stepy * double(j);
	double cxs[4] = {zx + hx, zx - hx, zx - hx, zx + hx};
	double cys[4] = {zy + hy, zy + hy, zy - hy, zy - hy};
	double yx[4], yy[4];
	bool bad[4];
	for (int k = 0; k < 4; k++)
	{
		Deflection d = deflect_reference(a, cxs[k], cys[k]);
		bad[k] = isnan(d.wx) || isnan(d.wy);
		yx[k] = ((d.wx - a.cyx) + a.h1) * a.n1 / (2 * a.h1);
		yy[k] = double(a.n2) - ((d.wy - a.cyy) + a.h2) * a.n2 / (2 * a.h2);
	}
	const int tris[2][3] = {{0, 1, 2}, {2, 3, 0}};
	for (int k = 0; k < 2; k++)
	{
		if (bad[tris[k][0]] || bad[tris[k][1]] || bad[tris[k][2]]) continue;
		refpoly::Poly P;
		P.n = 0;
		for (int v = 0; v < 3; v++) refpoly::add(P, yx[tris[k][v]], yy[tris[k][v]]);
		double A = refpoly::area(P);
		if (!(fabs(A) < 10000 * a.factor)) continue;
		double* map = reinterpret_cast<double*>(PARITY ? (A < 0 ? a.mmin : a.msad) : a.map);
		refpoly::allocate(P, a.factor, map, a.n1, a.n2);
	}
}

/* ---------------------------------------------------------------------------
IRS (irs_functions.cuh:111-212)
--------------------------------------------------------------------------- */
// irs_functions.cuh:152-184, array_functions.cuh:70-75, bit-for-bit the
// reference's operation order; -1 when the ray misses the map
__device__ __forceinline__ bool irs_pixel(const ShootArgs& a, double wx, double wy, int& px, int& py)
{
#pragma clang fp contract(off)
	double w1 = wx - a.cyx;
	double w2 = wy - a.cyy;
	if (w1 < -a.h1 || w1 > a.h1 || w2 < -a.h2 || w2 > a.h2) return false;
	px = int((w1 + a.h1) * a.n1 / (2 * a.h1));
	py = int((w2 + a.h2) * a.n2 / (2 * a.h2));
	if (px == a.n1) px--;
	if (py == a.n2) py--;
	py = a.n2 - 1 - py;
	return true;
}

__device__ __forceinline__ long long irs_bin(const ShootArgs& a, double wx, double wy)
{
	int px, py;
	if (!irs_pixel(a, wx, wy, px, py)) return -1;
	return (long long)py * a.n1 + px;
}

// One wave per block, each with its own LDS window: a block then waits for no slower wave
// before it flushes (10000^2 map: 0.261 s with four-wave blocks sharing a 128 x 64 window,
// 0.251 s with one-wave blocks and 64 x 32 windows; the same counts)
constexpr int IRS_THREADS = 64;
constexpr int IRS_TILE_H = 32;        // rows of rays a block shoots: the shortest tile
constexpr int IRS_TILE_H_MAX = 256;   // and the tallest

/* Rows per IRS tile: the tallest power of two in [IRS_TILE_H, IRS_TILE_H_MAX]
   that still gives eight blocks per resident block slot. A tall tile zeroes and
   flushes its LDS window once for more rays (10000^2 map: 0.323 s at 32 rows,
   0.304 s at 256). */
static inline int irs_tile_rows(long long ntx, long long rows, int num_cus)
{
	const long long want = 8LL * 3 * (num_cus > 0 ? num_cus : 256);
	int h = IRS_TILE_H_MAX;
	while (h > IRS_TILE_H && ntx * ((rows + h - 1) / h) < want) h >>= 1;
	return h;
}
constexpr int IRS_WX = 64, IRS_WY = 32;   // LDS window of ray counts, 8 KB: 16 waves use 128 KB of a CU

/*
Ray counts go through three levels: a per-lane two-slot register cache
(consecutive rows of a lane usually land in one pixel), a per-block LDS
window of IRS_WX x IRS_WY pixels centred on the block's first row of rays
(a block's rays land in a few dozen pixels away from stars), and the map in
HBM for everything else. Only evicted register slots touch LDS, only the
window's nonzero pixels and the rays that miss it touch HBM. Counts are
integers, so the map is independent of the order.
*/
struct IrsWindow
{
	int* w;
	int x0, y0;
};

__device__ __forceinline__ void irs_count(const ShootArgs& a, const IrsWindow& win, int px, int py, int v)
{
	const int dx = px - win.x0, dy = py - win.y0;
	if (dx >= 0 && dx < IRS_WX && dy >= 0 && dy < IRS_WY)
		atomicAdd(win.w + dy * IRS_WX + dx, v);
	else
		atomicAdd(reinterpret_cast<int*>(a.map) + ((long long)py * a.n1 + px), v);
}

template <int PREC, bool PARITY>
__global__ __launch_bounds__(IRS_THREADS) void k_irs_fast(ShootArgs a, long long row0, long long row1,
	unsigned long long* __restrict__ received, MLENS_TABLE_PARAMS, void* __restrict__ map, void* __restrict__ mmin,
	void* __restrict__ msad, int tile_h)
{
	MLENS_TABLE_REBIND(a)
	a.map = map;
	a.mmin = mmin;
	a.msad = msad;
	static_assert(IRS_THREADS == 64, "the window is placed from one wave's reduction of its first row");
	__shared__ int win_counts[IRS_WX * IRS_WY];
	const long long i = (long long)blockIdx.x * IRS_THREADS + threadIdx.x;
	const long long J0 = row0 + (long long)blockIdx.y * tile_h;
	if (J0 >= row1) return;
	const long long J1 = min(J0 + tile_h, row1);
	const bool ok = i < a.N1;
	const double zx = lattice_at(a.x0, a.stepx, ok ? i : a.N1 - 1);
	LeafCache cache;
	unsigned long long got = 0;
	for (int k = threadIdx.x; k < IRS_WX * IRS_WY; k += IRS_THREADS) win_counts[k] = 0;
	__syncthreads();

	// register cache: (x, y) of the two most recent pixels and their counts
	int kx0 = -1, ky0 = -1, c0 = 0, kx1 = -1, ky1 = -1, c1 = 0;
	PixelAcc<int> par;
	IrsWindow win{win_counts, 0, 0};
	for (long long j = J0; j < J1; j++)
	{
		double zy = lattice_at(a.y0, a.stepy, j);
		Deflection d = deflect<PREC, PARITY>(a, zx, zy, cache);
		int px = 0, py = 0;
		const bool hit = ok && !isnan(d.wx) && !isnan(d.wy) && irs_pixel(a, d.wx, d.wy, px, py);
		if (j == J0)
		{
			// window centred on the bounding box of the block's (the wave's) first row
			int mnx = hit ? px : 0x7fffffff, mny = hit ? py : 0x7fffffff;
			int mxx = hit ? px : -0x7fffffff, mxy = hit ? py : -0x7fffffff;
			for (int off = 32; off > 0; off >>= 1)
			{
				mnx = min(mnx, __shfl_xor(mnx, off));
				mny = min(mny, __shfl_xor(mny, off));
				mxx = max(mxx, __shfl_xor(mxx, off));
				mxy = max(mxy, __shfl_xor(mxy, off));
			}
			if (mnx <= mxx)
			{
				win.x0 = (mnx + mxx) / 2 - IRS_WX / 2;
				win.y0 = (mny + mxy) / 2 - IRS_WY / 2;
			}
			else
			{
				win.x0 = win.y0 = -0x3fffffff;   // no ray of the first row landed: no window
			}
		}
		if (!hit) continue;
		if (PARITY) acc_add<int, true>(a, par, ((long long)py * a.n1 + px) * 2 + (d.mu >= 0 ? 0 : 1), 1);
		if (px == kx0 && py == ky0)
		{
			c0++;
		}
		else if (px == kx1 && py == ky1)
		{
			const int t = c1 + 1;
			kx1 = kx0;
			ky1 = ky0;
			c1 = c0;
			kx0 = px;
			ky0 = py;
			c0 = t;
		}
		else
		{
			if (c1) irs_count(a, win, kx1, ky1, c1);
			kx1 = kx0;
			ky1 = ky0;
			c1 = c0;
			kx0 = px;
			ky0 = py;
			c0 = 1;
		}
		got++;
	}
	if (c0) irs_count(a, win, kx0, ky0, c0);
	if (c1) irs_count(a, win, kx1, ky1, c1);
	if (PARITY) acc_flush<int, true>(a, par);
	__syncthreads();
	for (int k = threadIdx.x; k < IRS_WX * IRS_WY; k += IRS_THREADS)
	{
		const int v = win_counts[k];
		if (v) atomicAdd(reinterpret_cast<int*>(a.map) + ((long long)(win.y0 + k / IRS_WX) * a.n1 + (win.x0 + k % IRS_WX)), v);
	}
	// wave reduction, one 64-bit atomic per wave
	for (int off = 32; off > 0; off >>= 1) got += __shfl_down(got, off);
	if ((threadIdx.x & 63) == 0 && got) atomicAdd(received, got);
}

template <int PREC>
__global__ void k_irs_positions(ShootArgs a, long long row, double* w, long long* bins)
{
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	bool ok = i < a.N1;
	double zx = lattice_at(a.x0, a.stepx, ok ? i : a.N1 - 1);
	double zy = lattice_at(a.y0, a.stepy, row);
	LeafCache cache;
	Deflection d = deflect<PREC, false>(a, zx, zy, cache);
	if (!ok) return;
	w[2 * i] = d.wx;
	w[2 * i + 1] = d.wy;
	bins[i] = (isnan(d.wx) || isnan(d.wy)) ? -2 : irs_bin(a, d.wx, d.wy);
}

template <int PREC, bool MU>
__global__ void k_deflect_points(ShootArgs a, const double* z, int n, double* w, double* mu)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	bool ok = i < n;
	double zx = ok ? z[2 * i] : 0, zy = ok ? z[2 * i + 1] : 0;
	LeafCache cache;
	Deflection d = deflect<PREC, MU>(a, zx, zy, cache);
	if (!ok) return;
	w[2 * i] = d.wx;
	w[2 * i + 1] = d.wy;
	if (MU) mu[i] = d.mu;
}

void Engine::alloc_maps(int n1, int n2, bool ipm, Precision prec, bool parities)
{
	npix_[0] = n1;
	npix_[1] = n2;
	map_int_ = !ipm;
	map_fp32_ = ipm && prec == PREC_MIXED;
	parities_ = parities;
	size_t elem = map_int_ ? sizeof(int) : (map_fp32_ ? sizeof(float) : sizeof(double));
	size_t bytes = size_t(n1) * n2 * elem;
	map_.alloc(bytes);
	if (parities)
	{
		map_min_.alloc(bytes);
		map_sad_.alloc(bytes);
	}
	zero_maps();
}

/*
Tiled accumulation maps for the fast IPM kernel: 128-byte tiles of
MapTile<MapT>::W x H pixels, ceil(n1 / W) tiles per row, and beside each an
span map of the same type and layout (span_add). Allocated zeroed on first
use (and again when the map's shape or type changes); fold_tiled_maps adds
them into the row-major maps and clears them, so they are zero between shots.
*/
void Engine::ensure_tiled_maps()
{
	const int W = map_fp32_ ? MapTile<float>::W : MapTile<double>::W, H = map_fp32_ ? MapTile<float>::H : MapTile<double>::H;
	const int tx = (npix_[0] + W - 1) / W;
	const size_t bytes = size_t(tx) * size_t((npix_[1] + H - 1) / H) * 128;
	const size_t sbytes = bytes;
	const bool have = parities_ ? (tmap_min_.ptr && span_min_.ptr && span_sad_.ptr) : span_.ptr != nullptr;
	if (tmap_.ptr && tiles_x_ == tx && tmap_bytes_ == bytes && have) return;
	tiles_x_ = tx;
	tmap_bytes_ = bytes;
	tmap_.alloc(bytes);
	MLENS_HIP_CHECK(hipMemsetAsync(tmap_.ptr, 0, bytes, stream_));
	tmap_min_.free();
	tmap_sad_.free();
	span_.free();
	span_min_.free();
	span_sad_.free();
	if (parities_)
	{
		tmap_min_.alloc(bytes);
		tmap_sad_.alloc(bytes);
		span_min_.alloc(sbytes);
		span_sad_.alloc(sbytes);
		MLENS_HIP_CHECK(hipMemsetAsync(tmap_min_.ptr, 0, bytes, stream_));
		MLENS_HIP_CHECK(hipMemsetAsync(tmap_sad_.ptr, 0, bytes, stream_));
		MLENS_HIP_CHECK(hipMemsetAsync(span_min_.ptr, 0, sbytes, stream_));
		MLENS_HIP_CHECK(hipMemsetAsync(span_sad_.ptr, 0, sbytes, stream_));
	}
	else
	{
		span_.alloc(sbytes);
		MLENS_HIP_CHECK(hipMemsetAsync(span_.ptr, 0, sbytes, stream_));
	}
}

// the fold of a whole map: a block of 4 waves takes 4 rows, one wave per row
template <typename MapT>
__global__ __launch_bounds__(256) void k_fold_rows(MapT* __restrict__ map, MapT* __restrict__ tiled, MapT* __restrict__ span,
	int n1, int n2, int tiles_x)
{
	const int y = blockIdx.x * 4 + (threadIdx.x >> 6);
	if (y >= n2) return;
	fold_row<MapT, false>(map + (long long)n1 * y, tiled, span, n1, tiles_x, y);
}

void Engine::fold_tiled_maps()
{
	if (!tmap_.ptr) return;
	const unsigned grid = unsigned((npix_[1] + 3) / 4);
	for (int k = 0; k < (parities_ ? 3 : 1); k++)
	{
		char* m = k == 0 ? map_.ptr : (k == 1 ? map_min_.ptr : map_sad_.ptr);
		char* t = k == 0 ? tmap_.ptr : (k == 1 ? tmap_min_.ptr : tmap_sad_.ptr);
		// with parities the kernel writes minima and saddles only, so the total has no spans
		char* sp = k == 0 ? (parities_ ? nullptr : span_.ptr) : (k == 1 ? span_min_.ptr : span_sad_.ptr);
		if (map_fp32_)
			k_fold_rows<float><<<grid, 256, 0, stream_>>>((float*)m, (float*)t, (float*)sp, npix_[0], npix_[1], tiles_x_);
		else
			k_fold_rows<double><<<grid, 256, 0, stream_>>>((double*)m, (double*)t, (double*)sp, npix_[0], npix_[1], tiles_x_);
	}
	MLENS_HIP_CHECK(hipGetLastError());
}

void Engine::zero_maps()
{
	size_t elem = map_int_ ? sizeof(int) : (map_fp32_ ? sizeof(float) : sizeof(double));
	size_t bytes = size_t(npix_[0]) * npix_[1] * elem;
	MLENS_HIP_CHECK(hipMemsetAsync(map_.ptr, 0, bytes, stream_));
	if (parities_)
	{
		MLENS_HIP_CHECK(hipMemsetAsync(map_min_.ptr, 0, bytes, stream_));
		MLENS_HIP_CHECK(hipMemsetAsync(map_sad_.ptr, 0, bytes, stream_));
	}
	// tiled and span maps are zero between shots unless a shot was cut short
	for (DeviceBuffer<char>* t : {&tmap_, &tmap_min_, &tmap_sad_, &span_, &span_min_, &span_sad_})
		if (t->ptr) MLENS_HIP_CHECK(hipMemsetAsync(t->ptr, 0, std::min(tmap_bytes_, t->count), stream_));
}

ShootStats Engine::shoot_ipm(const LensParams& lens, const GridParams& g, Precision prec, ShootKernel kern, long long row0,
	long long row1)
{
	// the kernels write FP32 maps in mixed precision and FP64 maps otherwise
	if (map_int_ || map_fp32_ != (prec == PREC_MIXED && kern == KERNEL_FAST))
		throw HipError("IPM map type does not match the requested precision", __FILE__, __LINE__);
	ShootArgs a = ShootArgsBuilder::make(lens, g, false);
	MLENS_TREE_ARGS(a)
	a.map = map_.ptr;
	a.mmin = parities_ ? map_min_.ptr : nullptr;
	a.msad = parities_ ? map_sad_.ptr : nullptr;
	if (row1 < 0 || row1 > a.N2) row1 = a.N2;
	if (row0 < 0) row0 = 0;
	ShootStats st;
	st.cells = (row1 > row0) ? (row1 - row0) * a.N1 : 0;
	if (st.cells == 0) return st;
	MLENS_HIP_CHECK(hipEventRecord(ev0_, stream_));
	if (kern == KERNEL_REFERENCE)
	{
		if (map_fp32_) throw HipError("the reference kernel accumulates an FP64 map", __FILE__, __LINE__);
		dim3 grid(unsigned((a.N1 + 15) / 16), unsigned((row1 - row0 + 15) / 16));
		dim3 block(16, 16);
		if (parities_)
			k_ipm_reference<true><<<grid, block, 0, stream_>>>(a, row0, row1);
		else
			k_ipm_reference<false><<<grid, block, 0, stream_>>>(a, row0, row1);
	}
	else
	{
		ensure_tiled_maps();
		a.tiles_x = tiles_x_;
		a.map = tmap_.ptr;
		a.mmin = parities_ ? tmap_min_.ptr : nullptr;
		a.msad = parities_ ? tmap_sad_.ptr : nullptr;
		a.span = parities_ ? span_min_.ptr : span_.ptr;
		a.span_sad = parities_ ? span_sad_.ptr : nullptr;
		const long long ntx = (a.N1 + IPM_TILE_W - 1) / IPM_TILE_W;
		const int th = tile_rows_ > 0 ? tile_rows_ : ipm_tile_rows(ntx, row1 - row0, num_cus_);
		last_tile_rows_ = th;
		if ((row1 - row0 + th - 1) / th > 65535) throw HipError("too many tiles of this height for one launch", __FILE__, __LINE__);
		dim3 grid(unsigned(ntx), unsigned((row1 - row0 + th - 1) / th));
		dim3 block(64 * IPM_WAVES);
		if (prec == PREC_FP64)
		{
			if (parities_) k_ipm_fast<PREC_FP64, true><<<grid, block, 0, stream_>>>(a, row0, row1, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad,
				a.span, a.span_sad, th);
			else k_ipm_fast<PREC_FP64, false><<<grid, block, 0, stream_>>>(a, row0, row1, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad,
				a.span, a.span_sad, th);
		}
		else
		{
			if (parities_) k_ipm_fast<PREC_MIXED, true><<<grid, block, 0, stream_>>>(a, row0, row1, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad,
				a.span, a.span_sad, th);
			else k_ipm_fast<PREC_MIXED, false><<<grid, block, 0, stream_>>>(a, row0, row1, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad,
				a.span, a.span_sad, th);
		}
		MLENS_HIP_CHECK(hipGetLastError());
		fold_tiled_maps();
	}
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipEventRecord(ev1_, stream_));
	MLENS_HIP_CHECK(hipEventSynchronize(ev1_));
	float ms = 0;
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
	st.ms = ms;
	return st;
}

ShootStats Engine::shoot_irs(const LensParams& lens, const GridParams& g, Precision prec, long long row0, long long row1)
{
	if (!map_int_) throw HipError("IRS needs an integer map", __FILE__, __LINE__);
	ShootArgs a = ShootArgsBuilder::make(lens, g, true);
	MLENS_TREE_ARGS(a)
	a.map = map_.ptr;
	a.mmin = parities_ ? map_min_.ptr : nullptr;
	a.msad = parities_ ? map_sad_.ptr : nullptr;
	if (row1 < 0 || row1 > a.N2) row1 = a.N2;
	if (row0 < 0) row0 = 0;
	ShootStats st;
	st.cells = (row1 > row0) ? (row1 - row0) * a.N1 : 0;
	if (st.cells == 0) return st;
	scratch_ll_.alloc(1);
	MLENS_HIP_CHECK(hipMemsetAsync(scratch_ll_.ptr, 0, sizeof(long long), stream_));
	unsigned long long* rec = reinterpret_cast<unsigned long long*>(scratch_ll_.ptr);
	const long long ntx = (a.N1 + IRS_THREADS - 1) / IRS_THREADS;
	const int th = tile_rows_ > 0 ? tile_rows_ : irs_tile_rows(ntx, row1 - row0, num_cus_);
	last_tile_rows_ = th;
	if ((row1 - row0 + th - 1) / th > 65535) throw HipError("too many tiles of this height for one launch", __FILE__, __LINE__);
	dim3 grid(unsigned(ntx), unsigned((row1 - row0 + th - 1) / th));
	MLENS_HIP_CHECK(hipEventRecord(ev0_, stream_));
	if (prec == PREC_FP64)
	{
		if (parities_) k_irs_fast<PREC_FP64, true><<<grid, IRS_THREADS, 0, stream_>>>(a, row0, row1, rec, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad, th);
		else k_irs_fast<PREC_FP64, false><<<grid, IRS_THREADS, 0, stream_>>>(a, row0, row1, rec, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad, th);
	}
	else
	{
		if (parities_) k_irs_fast<PREC_MIXED, true><<<grid, IRS_THREADS, 0, stream_>>>(a, row0, row1, rec, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad, th);
		else k_irs_fast<PREC_MIXED, false><<<grid, IRS_THREADS, 0, stream_>>>(a, row0, row1, rec, MLENS_TABLE_ARGS(a), a.map, a.mmin, a.msad, th);
	}
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipEventRecord(ev1_, stream_));
	MLENS_HIP_CHECK(hipEventSynchronize(ev1_));
	float ms = 0;
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
	st.ms = ms;
	long long got = 0;
	MLENS_HIP_CHECK(hipMemcpy(&got, scratch_ll_.ptr, sizeof(long long), hipMemcpyDeviceToHost));
	st.rays_received = got;
	return st;
}

void Engine::deflect_points(const LensParams& lens, const double* z, int n, Precision prec, double* w, double* mu)
{
	GridParams g;
	std::memset(&g, 0, sizeof(g));
	g.half_length_y[0] = g.half_length_y[1] = 1;
	g.num_pixels[0] = g.num_pixels[1] = 1;
	ShootArgs a = ShootArgsBuilder::make(lens, g, false);
	MLENS_TREE_ARGS(a)
	DeviceBuffer<double> dz, dw, dmu;
	dz.alloc(size_t(2) * n);
	dw.alloc(size_t(2) * n);
	dmu.alloc(n);
	MLENS_HIP_CHECK(hipMemcpy(dz.ptr, z, sizeof(double) * 2 * n, hipMemcpyHostToDevice));
	if (mu)
		k_deflect_points<PREC_FP64, true><<<blocks_for(n, 256), 256, 0, stream_>>>(a, dz.ptr, n, dw.ptr, dmu.ptr);
	else if (prec == PREC_FP64)
		k_deflect_points<PREC_FP64, false><<<blocks_for(n, 256), 256, 0, stream_>>>(a, dz.ptr, n, dw.ptr, nullptr);
	else
		k_deflect_points<PREC_MIXED, false><<<blocks_for(n, 256), 256, 0, stream_>>>(a, dz.ptr, n, dw.ptr, nullptr);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	MLENS_HIP_CHECK(hipMemcpy(w, dw.ptr, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
	if (mu) MLENS_HIP_CHECK(hipMemcpy(mu, dmu.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
}

/*
Deflection without allocation. One call's points, deflections and
magnifications are staged in `scratch`, a fixed array in the device's global
memory (2.6 MB) that every GPU running this library holds until the process
ends. Host code never uses the array's name as an address: the device address
is looked up once per device, with that device current, and kept. Calls that
overlap on one device share the array and race.
*/
static __device__ __align__(256) double scratch[5 * Engine::DEFLECT_NOALLOC_MAX];

constexpr int SCRATCH_MAX_DEVICES = 64;

// the device address of `scratch` on device dev, which must be current
static double* scratch_on(int dev)
{
	static std::mutex mtx;
	static double* addr[SCRATCH_MAX_DEVICES] = {};
	if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) throw HipError("device index beyond the scratch table", __FILE__, __LINE__);
	std::lock_guard<std::mutex> lock(mtx);
	if (!addr[dev])
	{
		void* p = nullptr;
		MLENS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		addr[dev] = static_cast<double*>(p);
	}
	return addr[dev];
}

void Engine::deflect_points_noalloc(const LensParams& lens, const double* z, int n, Precision prec, double* w, double* mu)
{
	if (n < 0 || n > DEFLECT_NOALLOC_MAX)
		throw HipError("ipm_deflect_noalloc takes 0 to " + std::to_string(DEFLECT_NOALLOC_MAX) + " points per call, not " +
				std::to_string(n),
			__FILE__, __LINE__);
	if (n == 0) return;
	int cur = 0;
	MLENS_HIP_CHECK(hipGetDevice(&cur));
	if (cur != dev_) MLENS_HIP_CHECK(hipSetDevice(dev_));
	struct Restore
	{
		int dev;
		bool on;
		~Restore()
		{
			if (on) (void)hipSetDevice(dev);
		}
	} restore{cur, cur != dev_};
	double* base = scratch_on(dev_);
	double* dz = base;
	double* dw = base + 2 * DEFLECT_NOALLOC_MAX;
	double* dmu = base + 4 * DEFLECT_NOALLOC_MAX;
	GridParams g;
	std::memset(&g, 0, sizeof(g));
	g.half_length_y[0] = g.half_length_y[1] = 1;
	g.num_pixels[0] = g.num_pixels[1] = 1;
	ShootArgs a = ShootArgsBuilder::make(lens, g, false);
	MLENS_TREE_ARGS(a)
	MLENS_HIP_CHECK(hipMemcpyAsync(dz, z, sizeof(double) * 2 * n, hipMemcpyHostToDevice, stream_));
	if (mu)
		k_deflect_points<PREC_FP64, true><<<blocks_for(n, 256), 256, 0, stream_>>>(a, dz, n, dw, dmu);
	else if (prec == PREC_FP64)
		k_deflect_points<PREC_FP64, false><<<blocks_for(n, 256), 256, 0, stream_>>>(a, dz, n, dw, nullptr);
	else
		k_deflect_points<PREC_MIXED, false><<<blocks_for(n, 256), 256, 0, stream_>>>(a, dz, n, dw, nullptr);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipMemcpyAsync(w, dw, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, stream_));
	if (mu) MLENS_HIP_CHECK(hipMemcpyAsync(mu, dmu, sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::irs_positions(const LensParams& lens, const GridParams& g, Precision prec, long long row, double* w)
{
	ShootArgs a = ShootArgsBuilder::make(lens, g, true);
	MLENS_TREE_ARGS(a)
	long long n = a.N1;
	DeviceBuffer<double> dw;
	DeviceBuffer<long long> db;
	dw.alloc(size_t(2) * n);
	db.alloc(n);
	if (prec == PREC_FP64)
		k_irs_positions<PREC_FP64><<<blocks_for(n, 256), 256, 0, stream_>>>(a, row, dw.ptr, db.ptr);
	else
		k_irs_positions<PREC_MIXED><<<blocks_for(n, 256), 256, 0, stream_>>>(a, row, dw.ptr, db.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	// layout: w[3 i] = w1, w[3 i + 1] = w2, w[3 i + 2] = GPU bin (as double)
	std::vector<double> hw(size_t(2) * n);
	std::vector<long long> hb(n);
	MLENS_HIP_CHECK(hipMemcpy(hw.data(), dw.ptr, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
	MLENS_HIP_CHECK(hipMemcpy(hb.data(), db.ptr, sizeof(long long) * n, hipMemcpyDeviceToHost));
	for (long long i = 0; i < n; i++)
	{
		w[3 * i] = hw[2 * i];
		w[3 * i + 1] = hw[2 * i + 1];
		w[3 * i + 2] = double(hb[i]);
	}
}

double Engine::mean_near_stars(const GridParams& g) const
{
	// average near-field count over the lattice, weighting each leaf by the
	// area it shares with the shooting rectangle
	std::vector<double> geom(size_t(num_leaves_) * 4);
	std::vector<int> nn(num_leaves_);
	MLENS_HIP_CHECK(hipMemcpy(geom.data(), leaf_geom_.ptr, geom.size() * sizeof(double), hipMemcpyDeviceToHost));
	MLENS_HIP_CHECK(hipMemcpy(nn.data(), leaf_near_true_.ptr, nn.size() * sizeof(int), hipMemcpyDeviceToHost));
	double x0 = g.center_x[0] - g.half_length_x[0], x1 = g.center_x[0] + g.half_length_x[0];
	double y0 = g.center_x[1] - g.half_length_x[1], y1 = g.center_x[1] + g.half_length_x[1];
	double num = 0, den = 0;
	for (int L = 0; L < num_leaves_; L++)
	{
		double cx = geom[4 * L], cy = geom[4 * L + 1], h = geom[4 * L + 2];
		double ox = std::max(0.0, std::min(x1, cx + h) - std::max(x0, cx - h));
		double oy = std::max(0.0, std::min(y1, cy + h) - std::max(y0, cy - h));
		num += ox * oy * nn[L];
		den += ox * oy;
	}
	return den > 0 ? num / den : 0;
}

/* ---------------------------------------------------------------------------
Map utilities (array_functions.cuh:86-160)
--------------------------------------------------------------------------- */
template <typename V>
__global__ void k_add3(V* total, const V* a, const V* b, long long n)
{
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) total[i] = a[i] + b[i];
}

void Engine::add_parity_maps()
{
	if (!parities_) return;
	long long n = (long long)npix_[0] * npix_[1];
	if (map_int_)
		k_add3<int><<<blocks_for(n, 256), 256, 0, stream_>>>((int*)map_.ptr, (int*)map_min_.ptr, (int*)map_sad_.ptr, n);
	else if (map_fp32_)
		k_add3<float><<<blocks_for(n, 256), 256, 0, stream_>>>((float*)map_.ptr, (float*)map_min_.ptr, (float*)map_sad_.ptr, n);
	else
		k_add3<double><<<blocks_for(n, 256), 256, 0, stream_>>>((double*)map_.ptr, (double*)map_min_.ptr, (double*)map_sad_.ptr, n);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::download_map(int which, void* host) const
{
	const DeviceBuffer<char>& b = which == 0 ? map_ : (which == 1 ? map_min_ : map_sad_);
	size_t elem = map_int_ ? sizeof(int) : (map_fp32_ ? sizeof(float) : sizeof(double));
	MLENS_HIP_CHECK(hipMemcpy(host, b.ptr, size_t(npix_[0]) * npix_[1] * elem, hipMemcpyDeviceToHost));
}

template <typename V>
__global__ void k_minmax(const V* v, long long n, double* out)
{
	__shared__ double slo[256], shi[256];
	double lo = 1e308, hi = -1e308;
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
	{
		double x = double(v[i]);
		lo = fmin(lo, x);
		hi = fmax(hi, x);
	}
	slo[threadIdx.x] = lo;
	shi[threadIdx.x] = hi;
	__syncthreads();
	for (int off = blockDim.x / 2; off > 0; off >>= 1)
	{
		if ((int)threadIdx.x < off)
		{
			slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + off]);
			shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + off]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0)
	{
		out[2 * blockIdx.x] = slo[0];
		out[2 * blockIdx.x + 1] = shi[0];
	}
}

void Engine::map_minmax(int which, double* mn, double* mx)
{
	const DeviceBuffer<char>& b = which == 0 ? map_ : (which == 1 ? map_min_ : map_sad_);
	long long n = (long long)npix_[0] * npix_[1];
	const int nb = 1024;
	scratch_d_.alloc(2 * nb);
	if (map_int_)
		k_minmax<int><<<nb, 256, 0, stream_>>>((const int*)b.ptr, n, scratch_d_.ptr);
	else if (map_fp32_)
		k_minmax<float><<<nb, 256, 0, stream_>>>((const float*)b.ptr, n, scratch_d_.ptr);
	else
		k_minmax<double><<<nb, 256, 0, stream_>>>((const double*)b.ptr, n, scratch_d_.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	std::vector<double> h(2 * nb);
	MLENS_HIP_CHECK(hipMemcpyAsync(h.data(), scratch_d_.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	double lo = 1e308, hi = -1e308;
	for (int i = 0; i < nb; i++)
	{
		lo = std::min(lo, h[2 * i]);
		hi = std::max(hi, h[2 * i + 1]);
	}
	*mn = lo;
	*mx = hi;
}

// array_functions.cuh:114-160: index = round(value * factor - min), or of log10(value)
template <typename V>
__global__ void k_histogram(const V* v, long long n, int log_mode, int factor, long long hmin, long long len,
	unsigned long long* hist)
{
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
	{
		double x = double(v[i]);
		long long idx;
		if (log_mode)
		{
			if (!(x > 0)) continue;
			idx = llround(log10(x) * factor - double(hmin));
		}
		else
		{
			idx = llround(1.0 * x * factor - double(hmin));
		}
		if (idx >= 0 && idx < len) atomicAdd(&hist[idx], 1ull);
	}
}

void Engine::map_histogram(int which, bool log_mode, int factor, long long hmin, long long len, long long* host_hist)
{
	const DeviceBuffer<char>& b = which == 0 ? map_ : (which == 1 ? map_min_ : map_sad_);
	long long n = (long long)npix_[0] * npix_[1];
	DeviceBuffer<long long> dh;
	dh.alloc(std::max(len, 1LL));
	MLENS_HIP_CHECK(hipMemsetAsync(dh.ptr, 0, sizeof(long long) * std::max(len, 1LL), stream_));
	unsigned long long* hp = reinterpret_cast<unsigned long long*>(dh.ptr);
	const int nb = 2048;
	if (map_int_)
		k_histogram<int><<<nb, 256, 0, stream_>>>((const int*)b.ptr, n, log_mode, factor, hmin, len, hp);
	else if (map_fp32_)
		k_histogram<float><<<nb, 256, 0, stream_>>>((const float*)b.ptr, n, log_mode, factor, hmin, len, hp);
	else
		k_histogram<double><<<nb, 256, 0, stream_>>>((const double*)b.ptr, n, log_mode, factor, hmin, len, hp);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	MLENS_HIP_CHECK(hipMemcpy(host_hist, dh.ptr, sizeof(long long) * len, hipMemcpyDeviceToHost));
}

}  // namespace mlens

namespace mlens
{

size_t Engine::batch_args_size() { return sizeof(ShootArgs); }

void Engine::fill_batch_args(void* out, const LensParams& lens, const GridParams& g) const
{
	if (!map_fp32_ || map_int_ || !map_.ptr)
		throw HipError("a batched instance needs its FP32 IPM map allocated", __FILE__, __LINE__);
	ShootArgs a = ShootArgsBuilder::make(lens, g, false);
	MLENS_TREE_ARGS(a)
	a.map = map_.ptr;
	a.mmin = nullptr;
	a.msad = nullptr;
	std::memcpy(out, &a, sizeof(a));
}

struct BatchShooter::Impl
{
	std::vector<Engine*> engines;
	std::vector<ShootArgs> args;
	std::vector<BatchMap> maps;
	std::vector<long long> tiles_x;   // per instance: column tiles
	std::vector<std::unique_ptr<DeviceBuffer<char>>> tiled;   // per instance: tiled accumulation and span maps
	bool dirty = true;
	int tile_h = IPM_TILE_H, bpc = 12, ncu = 0;
	double tune_ms = 0;
	bool tuned = false;
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	int graph_tile_h = -1, graph_bpc = -1;
	DeviceBuffer<char> d_args, d_maps;
	DeviceBuffer<long long> d_start;
	DeviceBuffer<unsigned long long> d_counter;
	DeviceBuffer<int> d_hist;
	long long total = 0;
	long long max_npix = 0;
	int max_rows = 0;

	long long tiles(int h) const
	{
		long long t = 0;
		for (size_t i = 0; i < args.size(); i++) t += tiles_x[i] * ((args[i].N2 + h - 1) / h);
		return t;
	}

	void upload_start(int h)
	{
		std::vector<long long> st(args.size() + 1, 0);
		for (size_t i = 0; i < args.size(); i++) st[i + 1] = st[i] + tiles_x[i] * ((args[i].N2 + h - 1) / h);
		total = st.back();
		d_start.alloc(st.size());
		MLENS_HIP_CHECK(hipMemcpy(d_start.ptr, st.data(), st.size() * sizeof(long long), hipMemcpyHostToDevice));
	}

	void upload()
	{
		const size_t n = args.size();
		d_args.alloc(n * sizeof(ShootArgs));
		MLENS_HIP_CHECK(hipMemcpy(d_args.ptr, args.data(), n * sizeof(ShootArgs), hipMemcpyHostToDevice));
		d_maps.alloc(n * sizeof(BatchMap));
		MLENS_HIP_CHECK(hipMemcpy(d_maps.ptr, maps.data(), n * sizeof(BatchMap), hipMemcpyHostToDevice));
		d_counter.alloc(1);
		d_hist.alloc(n * BATCH_HIST_BINS);
		dirty = false;
	}

	// zero kernel launch on the batch stream (graph-safe: no memset nodes)
	void launch_zero()
	{
		const unsigned bx = unsigned(std::min<long long>(1024, std::max<long long>(1, (max_npix + 255) / 256)));
		dim3 grid(bx, unsigned(args.size()));
		k_batch_zero<<<grid, 256, 0, stream>>>(reinterpret_cast<const BatchMap*>(d_maps.ptr),
			reinterpret_cast<unsigned*>(d_hist.ptr), d_counter.ptr);
		MLENS_HIP_CHECK(hipGetLastError());
	}

	void launch_shoot(int h, int blocks_per_cu, long long stride)
	{
		const unsigned blocks = unsigned(ncu * blocks_per_cu);
		k_ipm_batch<<<blocks, 64 * IPM_BATCH_WAVES, 0, stream>>>(reinterpret_cast<const ShootArgs*>(d_args.ptr), d_start.ptr,
			int(args.size()), h, total, stride, d_counter.ptr);
		MLENS_HIP_CHECK(hipGetLastError());
	}

	void launch_fold()
	{
		const unsigned bx = unsigned(std::min(1024, std::max(1, (max_rows + 3) / 4)));
		dim3 grid(bx, unsigned(args.size()));
		k_batch_fold<<<grid, 256, 0, stream>>>(reinterpret_cast<const BatchMap*>(d_maps.ptr));
		MLENS_HIP_CHECK(hipGetLastError());
	}

	void launch_hist()
	{
		const unsigned bx = unsigned(std::min<long long>(256, std::max<long long>(1, (max_npix + 1023) / 1024)));
		dim3 grid(bx, unsigned(args.size()));
		k_batch_hist<<<grid, 256, 0, stream>>>(reinterpret_cast<const BatchMap*>(d_maps.ptr),
			reinterpret_cast<unsigned*>(d_hist.ptr));
		MLENS_HIP_CHECK(hipGetLastError());
	}

	float elapsed()
	{
		float ms = 0;
		MLENS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
		return ms;
	}

	void autotune(int verbose)
	{
		// about an eighth of the tiles, spread over every instance
		const long long stride = 8;
		const int heights[3] = {32, 64, 128};
		const int resident[3] = {8, 10, 12};   // one-wave blocks; 12 fill a CU at 3 waves per SIMD
		float best = -1;
		auto t0 = std::chrono::steady_clock::now();
		for (int h : heights)
		{
			upload_start(h);
			for (int b : resident)
			{
				launch_zero();
				MLENS_HIP_CHECK(hipEventRecord(e0, stream));
				launch_shoot(h, b, stride);
				MLENS_HIP_CHECK(hipEventRecord(e1, stream));
				MLENS_HIP_CHECK(hipEventSynchronize(e1));
				const float ms = elapsed();
				if (verbose >= 2) std::cout << "batch tune: tile_h " << h << ", blocks/CU " << b << ": " << ms << " ms\n";
				if (best < 0 || ms < best)
				{
					best = ms;
					tile_h = h;
					bpc = b;
				}
			}
		}
		tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		tuned = true;
	}

	void capture()
	{
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
		exec = nullptr;
		graph = nullptr;
		upload_start(tile_h);
		MLENS_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
		launch_zero();
		launch_shoot(tile_h, bpc, 1);
		launch_fold();
		launch_hist();
		MLENS_HIP_CHECK(hipStreamEndCapture(stream, &graph));
		MLENS_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
		graph_tile_h = tile_h;
		graph_bpc = bpc;
	}
};

BatchShooter::BatchShooter() : impl_(new Impl)
{
	MLENS_HIP_CHECK(hipStreamCreateWithFlags(&impl_->stream, hipStreamNonBlocking));
	MLENS_HIP_CHECK(hipEventCreate(&impl_->e0));
	MLENS_HIP_CHECK(hipEventCreate(&impl_->e1));
	int dev = 0;
	MLENS_HIP_CHECK(hipGetDevice(&dev));
	MLENS_HIP_CHECK(hipDeviceGetAttribute(&impl_->ncu, hipDeviceAttributeMultiprocessorCount, dev));
}

BatchShooter::~BatchShooter()
{
	if (impl_->stream) (void)hipStreamSynchronize(impl_->stream);
	if (impl_->exec) (void)hipGraphExecDestroy(impl_->exec);
	if (impl_->graph) (void)hipGraphDestroy(impl_->graph);
	if (impl_->e0) (void)hipEventDestroy(impl_->e0);
	if (impl_->e1) (void)hipEventDestroy(impl_->e1);
	if (impl_->stream) (void)hipStreamDestroy(impl_->stream);
	delete impl_;
}

int BatchShooter::add(Engine* e, const LensParams& lens, const GridParams& g)
{
	ShootArgs a;
	e->fill_batch_args(&a, lens, g);
	if (a.N1 <= 0 || a.N2 <= 0) throw HipError("a batched instance has an empty lattice", __FILE__, __LINE__);
	// the kernel accumulates into a tiled map and a span map of its own; the fold writes the row-major map
	const int W = MapTile<float>::W, H = MapTile<float>::H;
	const int tx = (a.n1 + W - 1) / W;
	const size_t tbytes = size_t(tx) * size_t((a.n2 + H - 1) / H) * 128;
	const size_t sbytes = tbytes;
	auto buf = std::make_unique<DeviceBuffer<char>>();
	buf->alloc(tbytes);
	MLENS_HIP_CHECK(hipMemset(buf->ptr, 0, tbytes));
	auto sbuf = std::make_unique<DeviceBuffer<char>>();
	sbuf->alloc(sbytes);
	MLENS_HIP_CHECK(hipMemset(sbuf->ptr, 0, sbytes));
	BatchMap bm{reinterpret_cast<float*>(a.map), reinterpret_cast<float*>(buf->ptr), reinterpret_cast<float*>(sbuf->ptr),
		(long long)a.n1 * a.n2, (long long)(tbytes / sizeof(float)), a.n1, a.n2, tx};
	a.map = buf->ptr;
	a.span = sbuf->ptr;
	a.span_sad = nullptr;
	a.tiles_x = tx;
	impl_->tiled.push_back(std::move(buf));
	impl_->tiled.push_back(std::move(sbuf));
	impl_->engines.push_back(e);
	impl_->args.push_back(a);
	impl_->maps.push_back(bm);
	impl_->tiles_x.push_back((a.N1 + IPM_BATCH_TILE_W - 1) / IPM_BATCH_TILE_W);
	impl_->max_npix = std::max(impl_->max_npix, (long long)a.n1 * a.n2);
	impl_->max_rows = std::max(impl_->max_rows, a.n2);
	impl_->dirty = true;
	impl_->graph_tile_h = -1;
	return int(impl_->args.size()) - 1;
}

double BatchShooter::shoot(int verbose)
{
	Impl& m = *impl_;
	if (m.args.empty()) return 0;
	// the instances' engines built their tables on their own streams
	for (Engine* e : m.engines) MLENS_HIP_CHECK(hipStreamSynchronize(e->stream()));
	if (m.dirty) m.upload();
	if (!m.tuned) m.autotune(verbose);
	if (m.graph_tile_h != m.tile_h || m.graph_bpc != m.bpc || !m.exec) m.capture();
	MLENS_HIP_CHECK(hipEventRecord(m.e0, m.stream));
	MLENS_HIP_CHECK(hipGraphLaunch(m.exec, m.stream));
	MLENS_HIP_CHECK(hipEventRecord(m.e1, m.stream));
	MLENS_HIP_CHECK(hipEventSynchronize(m.e1));
	return m.elapsed();
}

void BatchShooter::histograms(unsigned* host) const
{
	const size_t n = impl_->args.size() * BATCH_HIST_BINS;
	if (n) MLENS_HIP_CHECK(hipMemcpy(host, impl_->d_hist.ptr, n * sizeof(unsigned), hipMemcpyDeviceToHost));
}

int BatchShooter::size() const { return int(impl_->args.size()); }
long long BatchShooter::cells() const
{
	long long c = 0;
	for (const ShootArgs& a : impl_->args) c += a.N1 * a.N2;
	return c;
}
int BatchShooter::tile_h() const { return impl_->tile_h; }
int BatchShooter::blocks_per_cu() const { return impl_->bpc; }
double BatchShooter::tune_ms() const { return impl_->tune_ms; }
void BatchShooter::set_tuning(int tile_h, int blocks_per_cu)
{
	if (tile_h < 1 || tile_h > 4096 || blocks_per_cu < 1 || blocks_per_cu > 16)
		throw HipError("batch tuning outside 1..4096 rows and 1..16 blocks per CU", __FILE__, __LINE__);
	impl_->tile_h = tile_h;
	impl_->bpc = blocks_per_cu;
	impl_->tuned = true;
}

}  // namespace mlens
