/*
Number of caustic crossings (ncc_functions.cuh:14-411, ncc.cuh:269-330 of
weisluke/irs) on the MI355X.

The reference gives each caustic segment a thread that, for every pixel row
whose centre the segment crosses, adds +1 or -1 to every pixel left of the
crossing: one global atomic per pixel, O(crossings x width). Here a crossing
adds its +1 or -1 once, at the crossing's column of a row-difference grid;
suffix sums along each row then give exactly the reference's counts (integer
sums do not depend on their order). The over-sampled grid is reduced to the
final one by the maximum over each 2^over_sample x 2^over_sample block, which
is what the reference's repeated 2 x 2 reductions and shifts
(ncc_functions.cuh:345-411) leave behind; one workgroup per final row does
the sums and the maxima together, so the over-sampled counts never go back to
memory.

Segment clipping and crossing positions follow ncc_functions.cuh:43-301 step
by step in FP64. The reference evaluates the line intersection as
exp(log|y - y1| - log|dy| + log|dx|) with the signs restored; here it is the
direct product (y - y1) dx / dy, which differs from it only in the last
bits. A horizontal line asked for its x (or a vertical one for its y) gives
NaN, as the reference's 0 x inf does. Floating-point contraction is off in
this file, so the FP64 oracle's plain IEEE operations reproduce every
crossing bit for bit.
*/
#include "ncc.hpp"
#include "wave_scan.hpp"

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <stdexcept>
#include <string>

#pragma clang fp contract(off)

namespace mlens
{
namespace
{

struct NCCArgs
{
	double cx, cy;     // centre of the map
	double hx, hy;     // half-lengths
	int npx, npy;      // fine pixels per axis (final pixels << over_sample)
};

__device__ __forceinline__ double ncc_sgn(double v) { return v < 0 ? -1.0 : (v > 0 ? 1.0 : 0.0); }

// conversion to int as the GPU's cvt does it: NaN -> 0, out of range -> saturated
__device__ __forceinline__ long long ncc_int(double v)
{
	if (v != v) return 0;
	if (v >= 2147483647.0) return INT_MAX;
	if (v <= -2147483648.0) return INT_MIN;
	return (long long)(int)v;
}

// x of the line through p1, p2 at height y (get_x_intersection, ncc_functions.cuh:58-84)
__device__ __forceinline__ double line_x_at(double y, double p1x, double p1y, double p2x, double p2y)
{
	const double dx = p2x - p1x;
	if (dx == 0) return p1x;
	const double dy = p2y - p1y;
	if (dy == 0) return __builtin_nan("");
	return p1x + (y - p1y) * dx / dy;
}

// y of the line through p1, p2 at abscissa x (get_y_intersection, ncc_functions.cuh:85-110)
__device__ __forceinline__ double line_y_at(double x, double p1x, double p1y, double p2x, double p2y)
{
	const double dy = p2y - p1y;
	if (dy == 0) return p1y;
	const double dx = p2x - p1x;
	if (dx == 0) return __builtin_nan("");
	return p1y + (x - p1x) * dy / dx;
}

// the half-strip x > -hx, |y| < hy that a ray to +infinity in x from a pixel
// centre can meet (point_in_region, ncc_functions.cuh:43-50)
__device__ __forceinline__ bool in_region(double x, double y, const NCCArgs& g) { return x > -g.hx && fabs(y) < g.hy; }

// moves p1 along the line to p2 to a hundredth of a fine pixel inside the
// region (corrected_point, ncc_functions.cuh:124-155)
__device__ __forceinline__ void corrected(double p1x, double p1y, double p2x, double p2y, const NCCArgs& g, double& x,
	double& y)
{
	x = p1x;
	y = p1y;
	if (x <= -g.hx)
	{
		x = -g.hx + 0.01 * 2 * g.hx / double(g.npx);
		y = line_y_at(x, p1x, p1y, p2x, p2y);
	}
	if (fabs(y) >= g.hy)
	{
		y = ncc_sgn(y) * (g.hy - 0.01 * 2 * g.hy / double(g.npy));
		x = line_x_at(y, p1x, p1y, p2x, p2y);
	}
}

/*
The crossings (find_num_caustic_crossings_kernel, ncc_functions.cuh:170-301).
A row crossed upwards (the caustics are traced clockwise) takes 1 from every
pixel left of the crossing, a row crossed downwards adds 1: here at the
crossing's column of the difference grid. Each lane clips one segment and
counts its crossings; the wave then deals all crossings of its 64 segments
out to its lanes, 64 at a time, so that a long segment near a cusp (hundreds
of rows) does not hold up the wave: a lane finds the segment of crossing c by
a binary search over the lanes' first crossing indices. counters[0] counts the
crossings, counters[1] those whose row is outside the grid (dropped, as the
reference's bounds check drops them). One wave per workgroup.
*/
__global__ __launch_bounds__(64) void k_ncc_crossings(const double* __restrict__ caus, int nrows, int ncols, NCCArgs g,
	int* __restrict__ diff, unsigned long long* __restrict__ counters)
{
	const int lane = threadIdx.x & 63;
	const long long nseg = (long long)nrows * (ncols - 1);
	const long long stride = (long long)gridDim.x * blockDim.x;
	unsigned long long made = 0, dropped = 0;
	for (long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x - lane); base < nseg; base += stride)
	{
		const long long sg = base + lane;
		long long n = 0, k0 = 0;
		double px0 = 0, py0 = 0, px1 = 0, py1 = 0;
		int up = 0;
		if (sg < nseg)
		{
			const long long i = sg / (ncols - 1), j = sg - i * (ncols - 1);
			const double* p = caus + 2 * (i * ncols + j);
			const double x0 = p[0] - g.cx, y0 = p[1] - g.cy, x1 = p[2] - g.cx, y1 = p[3] - g.cy;
			bool take = in_region(x0, y0, g) || in_region(x1, y1, g);
			if (!take)
				take = line_x_at(g.hy, x0, y0, x1, y1) >= -g.hx || line_x_at(-g.hy, x0, y0, x1, y1) >= -g.hx ||
					fabs(line_y_at(-g.hx, x0, y0, x1, y1)) <= g.hy;
			if (take)
			{
				double ax, ay, bx, by;
				corrected(x0, y0, x1, y1, g, ax, ay);
				corrected(x1, y1, ax, ay, g, bx, by);
				// point_to_pixel (array_functions.cuh:70-75)
				px0 = (ax + g.hx) * double(g.npx) / (2 * g.hx);
				py0 = (ay + g.hy) * double(g.npy) / (2 * g.hy);
				px1 = (bx + g.hx) * double(g.npx) / (2 * g.hx);
				py1 = (by + g.hy) * double(g.npy) / (2 * g.hy);
				// up: centres k + 0.5 for k in [k0, k1); down: centres k - 0.5 for k in (k1, k0]
				k0 = ncc_int(py0 + 0.5);
				const long long k1 = ncc_int(py1 + 0.5);
				if (py0 < py1)
				{
					up = 1;
					n = k1 > k0 ? k1 - k0 : 0;
				}
				else if (py0 > py1)
				{
					n = k0 > k1 ? k0 - k1 : 0;
				}
			}
		}
		// crossings of lanes 0..lane, and the wave's total
		long long incl = n;
		for (int d = 1; d < 64; d <<= 1)
		{
			const long long u = __shfl_up(incl, d);
			if (lane >= d) incl += u;
		}
		const long long total = __shfl(incl, 63);
		const long long first = incl - n;
		for (long long c0 = 0; c0 < total; c0 += 64)
		{
			const long long c = c0 + lane;
			// the last lane whose first crossing is at or before c owns crossing c
			int o = 0;
			for (int step = 32; step > 0; step >>= 1)
				if (__shfl(first, o + step) <= c) o += step;
			const long long t = c - __shfl(first, o);
			const double qx0 = __shfl(px0, o), qy0 = __shfl(py0, o), qx1 = __shfl(px1, o), qy1 = __shfl(py1, o);
			const long long q0 = __shfl(k0, o);
			const bool qup = __shfl(up, o) != 0;
			if (c >= total) continue;
			const long long k = qup ? q0 + t : q0 - t;
			const double yc = qup ? double(k) + 0.5 : double(k) - 0.5;
			const double xc = line_x_at(yc, qx0, qy0, qx1, qy1);
			long long col = ncc_int(xc + 0.5) - 1;
			const long long row = qup ? (long long)g.npy - 1 - k : (long long)g.npy - k;
			if (col > g.npx - 1) col = g.npx - 1;
			if (col < 0) continue;
			if (row < 0 || row >= g.npy)
			{
				dropped++;
				continue;
			}
			atomicAdd(diff + row * g.npx + col, qup ? -1 : 1);
			made++;
		}
	}
	// one pair of counter atomics per wave
	for (int o = 32; o > 0; o >>= 1)
	{
		made += __shfl_xor(made, o);
		dropped += __shfl_xor(dropped, o);
	}
	if (lane == 0 && (made | dropped))
	{
		atomicAdd(counters, made);
		atomicAdd(counters + 1, dropped);
	}
}

/*
Each fine row, in place: count[x] = sum of diff[c] over c >= x. Tiles of
256 columns from the right; thread t takes column end - 1 - t, so an
inclusive scan over the threads sums the tile from the right.
*/
__global__ __launch_bounds__(256) void k_ncc_row_sums(int* __restrict__ fine, int width)
{
	__shared__ int wsum[4];
	int* r = fine + (long long)blockIdx.x * width;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int carry = 0;
	for (int end = width; end > 0; end -= 256)
	{
		const int col = end - 1 - int(threadIdx.x);
		const int v = wave_prefix_sum(col >= 0 ? r[col] : 0);
		if (lane == 63) wsum[wave] = v;
		__syncthreads();
		int before = carry;
		for (int w = 0; w < wave; w++) before += wsum[w];
		if (col >= 0) r[col] = v + before;
		carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
		__syncthreads();
	}
}

// final pixel (i, j) = maximum over its s x s block of fine pixels
__global__ __launch_bounds__(256) void k_ncc_block_max(const int* __restrict__ fine, int* __restrict__ map, int np1, int np2,
	int s)
{
	const long long n = (long long)np1 * np2;
	const long long width = (long long)np1 * s;
	for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
	{
		const long long j = q / np1, i = q - j * np1;
		const int* b = fine + j * s * width + i * s;
		int m = INT_MIN;
		for (int a = 0; a < s; a++)
			for (int c = 0; c < s; c++) m = max(m, b[a * width + c]);
		map[q] = m;
	}
}

/*
The row sums and the block maxima in one pass, for final rows of at most
NCC_FUSED_COLS pixels: workgroup j owns final row j and keeps the running
maximum of each of its pixels in LDS; its waves take the s fine rows of that
row in turn, each scanning its row from the right in tiles of 64 columns with
DPP scans (four tiles loaded ahead), so the fine grid is read once and never
written back. Lane l of a tile holds column end - 1 - l, and a tile starts at
a multiple of min(s, 64), so g = min(s, 64) consecutive lanes hold columns of
one final pixel: they take their maximum by DPP and one of them merges it
into LDS.
*/
constexpr int NCC_FUSED_COLS = 8192;

__global__ __launch_bounds__(256) void k_ncc_rows_to_map(const int* __restrict__ diff, int* __restrict__ map, int np1,
	int s, int* __restrict__ row_ext)
{
	__shared__ int cmax[NCC_FUSED_COLS];
	__shared__ int part[2][4];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int width = np1 * s;
	const int g = s < 64 ? s : 64;
	for (int c = threadIdx.x; c < np1; c += blockDim.x) cmax[c] = INT_MIN;
	__syncthreads();
	for (int a = wave; a < s; a += int(blockDim.x >> 6))
	{
		const int* r = diff + ((long long)blockIdx.x * s + a) * width;
		int carry = 0;
		for (int end = width; end > 0; end -= 256)
		{
			int v[4];
#pragma unroll
			for (int t = 0; t < 4; t++)
			{
				const int col = end - 1 - 64 * t - lane;
				v[t] = col >= 0 ? r[col] : 0;
			}
#pragma unroll
			for (int t = 0; t < 4; t++)
			{
				const int col = end - 1 - 64 * t - lane;
				const int x = wave_prefix_sum(v[t]) + carry;
				carry = __builtin_amdgcn_readlane(x, 63);
				const int m = wave_group_max(col >= 0 ? x : INT_MIN, g);
				if (col >= 0 && (lane & (g - 1)) == 0) atomicMax(cmax + col / s, m);
			}
		}
	}
	__syncthreads();
	// the final row, and its minimum and maximum for the map's extremes
	int* out = map + (long long)blockIdx.x * np1;
	int lo = INT_MAX, hi = INT_MIN;
	for (int c = threadIdx.x; c < np1; c += blockDim.x)
	{
		const int v = cmax[c];
		out[c] = v;
		lo = min(lo, v);
		hi = max(hi, v);
	}
	for (int o = 32; o > 0; o >>= 1)
	{
		lo = min(lo, __shfl_xor(lo, o));
		hi = max(hi, __shfl_xor(hi, o));
	}
	if (lane == 0)
	{
		part[0][wave] = lo;
		part[1][wave] = hi;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		for (int w = 1; w < int(blockDim.x >> 6); w++)
		{
			lo = min(lo, part[0][w]);
			hi = max(hi, part[1][w]);
		}
		row_ext[2 * blockIdx.x] = lo;
		row_ext[2 * blockIdx.x + 1] = hi;
	}
}

// the map's extremes from the per-row ones: one workgroup
__global__ __launch_bounds__(256) void k_ncc_row_extremes(const int* __restrict__ row_ext, int rows, int* __restrict__ ext)
{
	__shared__ int part[2][4];
	int lo = INT_MAX, hi = INT_MIN;
	for (int q = threadIdx.x; q < rows; q += blockDim.x)
	{
		lo = min(lo, row_ext[2 * q]);
		hi = max(hi, row_ext[2 * q + 1]);
	}
	for (int o = 32; o > 0; o >>= 1)
	{
		lo = min(lo, __shfl_xor(lo, o));
		hi = max(hi, __shfl_xor(hi, o));
	}
	if ((threadIdx.x & 63) == 0)
	{
		part[0][threadIdx.x >> 6] = lo;
		part[1][threadIdx.x >> 6] = hi;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		for (int w = 1; w < int(blockDim.x >> 6); w++)
		{
			lo = min(lo, part[0][w]);
			hi = max(hi, part[1][w]);
		}
		ext[0] = lo;
		ext[1] = hi;
	}
}

__global__ void k_ncc_init(int* ext, unsigned long long* counters)
{
	if (threadIdx.x == 0)
	{
		ext[0] = INT_MAX;
		ext[1] = INT_MIN;
		counters[0] = 0;
		counters[1] = 0;
	}
}

// minimum and maximum of the map: one pair of atomics per workgroup (a few
// dozen workgroups), not per wave, so they do not queue on two addresses
__global__ __launch_bounds__(256) void k_ncc_extremes(const int* __restrict__ map, long long n, int* __restrict__ ext)
{
	__shared__ int part[2][4];
	int lo = INT_MAX, hi = INT_MIN;
	for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
	{
		lo = min(lo, map[q]);
		hi = max(hi, map[q]);
	}
	for (int o = 32; o > 0; o >>= 1)
	{
		lo = min(lo, __shfl_xor(lo, o));
		hi = max(hi, __shfl_xor(hi, o));
	}
	if ((threadIdx.x & 63) == 0)
	{
		part[0][threadIdx.x >> 6] = lo;
		part[1][threadIdx.x >> 6] = hi;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		for (int w = 1; w < int(blockDim.x >> 6); w++)
		{
			lo = min(lo, part[0][w]);
			hi = max(hi, part[1][w]);
		}
		atomicMin(ext, lo);
		atomicMax(ext + 1, hi);
	}
}

constexpr int NCC_LDS_BINS = 8192;

// histogram_kernel (array_functions.cuh:115-131) with integer values: bins of
// width 1 from vmin; per-workgroup counts in LDS when the range fits
template <bool LOCAL>
__global__ __launch_bounds__(256) void k_ncc_histogram(const int* __restrict__ map, long long n, int vmin, int len,
	unsigned long long* __restrict__ hist)
{
	__shared__ unsigned int local[LOCAL ? NCC_LDS_BINS : 1];
	if (LOCAL)
	{
		for (int b = threadIdx.x; b < len; b += blockDim.x) local[b] = 0;
		__syncthreads();
	}
	for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
	{
		const int b = map[q] - vmin;
		if (b < 0 || b >= len) continue;   // values outside [vmin, vmax] are not counted
		if (LOCAL)
			atomicAdd(local + b, 1u);
		else
			atomicAdd(hist + b, 1ull);
	}
	if (LOCAL)
	{
		__syncthreads();
		for (int b = threadIdx.x; b < len; b += blockDim.x)
			if (local[b]) atomicAdd(hist + b, (unsigned long long)local[b]);
	}
}

__global__ void k_ncc_zero(unsigned long long* v, int n)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q < n) v[q] = 0;
}

unsigned grid_for(long long n, int threads = 256, long long cap = 8192)
{
	const long long b = (n + threads - 1) / threads;
	return unsigned(b < 1 ? 1 : (b > cap ? cap : b));
}

// reductions that end in atomics on a few addresses: a few workgroups per CU
constexpr long long NCC_REDUCE_GRID = 64;

}  // namespace

NCCDevice::NCCDevice()
{
	MLENS_HIP_CHECK(hipStreamCreate(&stream_));
	MLENS_HIP_CHECK(hipEventCreate(&ev0_));
	MLENS_HIP_CHECK(hipEventCreate(&ev1_));
	MLENS_HIP_CHECK(hipEventCreate(&ev2_));
}

NCCDevice::~NCCDevice()
{
	if (ev0_) (void)hipEventDestroy(ev0_);
	if (ev1_) (void)hipEventDestroy(ev1_);
	if (ev2_) (void)hipEventDestroy(ev2_);
	if (stream_) (void)hipStreamDestroy(stream_);
}

void NCCDevice::count(const double* caustics, int nrows, int ncols, const NCCGrid& g, int* out, NCCStats& st)
{
	if (nrows < 1 || ncols < 2) throw std::runtime_error("the caustics need at least one row of two points");
	if (g.num_pixels_re < 1 || g.num_pixels_im < 1 || g.over_sample < 0)
		throw std::runtime_error("num_pixels_y1, num_pixels_y2 must be > 0 and over_sample >= 0");
	if (g.over_sample > 30) throw std::runtime_error("over_sample is too large for the pixel grid");
	const long long s = 1ll << g.over_sample;
	const long long npx = (long long)g.num_pixels_re * s, npy = (long long)g.num_pixels_im * s;
	if (npx > INT_MAX || npy > INT_MAX)
		throw std::runtime_error("num_pixels << over_sample must fit in a 32-bit int per axis");
	size_t free_b = 0, total_b = 0;
	MLENS_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
	const double need = double(npx) * double(npy) * sizeof(int);
	if (need > 0.9 * double(free_b))
		throw std::runtime_error("the over-sampled pixel grid (" + std::to_string((long long)(need / 1048576)) +
			" MiB) does not fit in device memory");

	const size_t npts = size_t(nrows) * size_t(ncols);
	if (caustics_.count < 2 * npts) caustics_.alloc(2 * npts);
	const size_t nfine = size_t(npx) * size_t(npy);
	if (fine_.count < nfine) fine_.alloc(nfine);
	const long long npix = (long long)g.num_pixels_re * g.num_pixels_im;
	if (map_.count < size_t(npix)) map_.alloc(size_t(npix));
	if (extremes_.count < 2) extremes_.alloc(2);
	if (counters_.count < 2) counters_.alloc(2);
	map_pixels_ = npix;

	MLENS_HIP_CHECK(hipMemcpyAsync(caustics_.ptr, caustics, 2 * npts * sizeof(double), hipMemcpyHostToDevice, stream_));
	MLENS_HIP_CHECK(hipMemsetAsync(fine_.ptr, 0, nfine * sizeof(int), stream_));
	k_ncc_init<<<1, 64, 0, stream_>>>(extremes_.ptr, counters_.ptr);
	MLENS_HIP_CHECK(hipGetLastError());

	NCCArgs a{g.center_re, g.center_im, g.half_length_re, g.half_length_im, int(npx), int(npy)};
	const long long nseg = (long long)nrows * (ncols - 1);
	MLENS_HIP_CHECK(hipEventRecord(ev0_, stream_));
	// one wave per workgroup: a few tens of thousands of segments still reach every CU
	k_ncc_crossings<<<grid_for(nseg, 64, 1 << 16), 64, 0, stream_>>>(caustics_.ptr, nrows, ncols, a, fine_.ptr, counters_.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipEventRecord(ev1_, stream_));
	const bool fused = g.num_pixels_re <= NCC_FUSED_COLS;
	if (fused)
	{
		if (row_ext_.count < size_t(2) * size_t(g.num_pixels_im)) row_ext_.alloc(size_t(2) * size_t(g.num_pixels_im));
		k_ncc_rows_to_map<<<unsigned(g.num_pixels_im), 256, 0, stream_>>>(fine_.ptr, map_.ptr, g.num_pixels_re, int(s),
			row_ext_.ptr);
		MLENS_HIP_CHECK(hipGetLastError());
	}
	else
	{
		k_ncc_row_sums<<<unsigned(npy), 256, 0, stream_>>>(fine_.ptr, int(npx));
		MLENS_HIP_CHECK(hipGetLastError());
		k_ncc_block_max<<<grid_for(npix), 256, 0, stream_>>>(fine_.ptr, map_.ptr, g.num_pixels_re, g.num_pixels_im, int(s));
		MLENS_HIP_CHECK(hipGetLastError());
	}
	MLENS_HIP_CHECK(hipEventRecord(ev2_, stream_));
	if (fused)
		k_ncc_row_extremes<<<1, 256, 0, stream_>>>(row_ext_.ptr, g.num_pixels_im, extremes_.ptr);
	else
		k_ncc_extremes<<<grid_for(npix, 256, NCC_REDUCE_GRID), 256, 0, stream_>>>(map_.ptr, npix, extremes_.ptr);
	MLENS_HIP_CHECK(hipGetLastError());

	MLENS_HIP_CHECK(hipMemcpyAsync(out, map_.ptr, size_t(npix) * sizeof(int), hipMemcpyDeviceToHost, stream_));
	int ext[2];
	unsigned long long cnt[2];
	MLENS_HIP_CHECK(hipMemcpyAsync(ext, extremes_.ptr, sizeof(ext), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipMemcpyAsync(cnt, counters_.ptr, sizeof(cnt), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	float ms = 0;
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
	st.t_ncc_ms = ms;
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, ev1_, ev2_));
	st.t_reduce_ms = ms;
	st.min = ext[0];
	st.max = ext[1];
	st.crossings = (long long)cnt[0];
	st.dropped = (long long)cnt[1];
}

void NCCDevice::histogram(int vmin, int vmax, long long* hist)
{
	if (map_pixels_ < 1) throw std::runtime_error("no map to take a histogram of");
	if (vmax < vmin) throw std::runtime_error("empty histogram range");
	const long long len = (long long)vmax - vmin + 1;
	if (len > INT_MAX) throw std::runtime_error("histogram range too large");
	if (hist_.count < size_t(len)) hist_.alloc(size_t(len));
	k_ncc_zero<<<unsigned((len + 255) / 256), 256, 0, stream_>>>(hist_.ptr, int(len));
	MLENS_HIP_CHECK(hipGetLastError());
	if (len <= NCC_LDS_BINS)
		k_ncc_histogram<true><<<grid_for(map_pixels_, 256, NCC_REDUCE_GRID), 256, 0, stream_>>>(map_.ptr, map_pixels_, vmin,
			int(len), hist_.ptr);
	else
		k_ncc_histogram<false><<<grid_for(map_pixels_, 256, NCC_REDUCE_GRID), 256, 0, stream_>>>(map_.ptr, map_pixels_, vmin,
			int(len), hist_.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	static_assert(sizeof(long long) == sizeof(unsigned long long), "histogram counts");
	MLENS_HIP_CHECK(hipMemcpyAsync(hist, hist_.ptr, size_t(len) * sizeof(long long), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
}

}  // namespace mlens
