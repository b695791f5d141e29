// Critical curves and caustics of a star field (ccf_functions.cuh, ccf.cuh).
//
// For every phase phi the critical curve points are the roots of
//   f(z, phi) = gamma - conj(d alpha / d zbar) - (1 - kappa - d alpha_smooth / dz) e^{-i phi},
// a rational function with a double pole at every star. All roots of one phi
// are refined together by Aberth's method (each Newton step corrected by the
// poles and by the other roots), starting at the centre phase of each branch
// from positions next to the stars, then stepping phi outwards in both
// directions with the previous phase's roots as starting points.
//
// On this GPU eight lanes update one root, a block of 256 lanes 32 roots of one
// phase. The sums over all roots and all stars of every update stream through
// LDS tiles, split over the root's eight lanes, so a root update costs
// O((roots + stars) / 8) arithmetic per lane with one global read per tile
// element per block. Updates are simultaneous (Jacobi): an iteration reads one
// copy of the roots and writes the other, so the result does not depend on
// the order in which blocks run.
#include "shoot_common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mlens
{

namespace
{

constexpr int CCF_THREADS = 256;
constexpr double CCF_PI = 3.14159265358979323846;

// the derivative terms of the lens equation at z, in the node z falls in
struct CCTerms
{
	dcx S1, S2;     // sum over the node's near stars of m / (z - z_i)^2 and m / (z - z_i)^3
	dcx L1, L2;     // theta^2 / h^2 sum i(i-1) c_i dz^(i-2) and theta^2 / h^3 sum i(i-1)(i-2) c_i dz^(i-3)
	double ddz;     // d alpha_smooth / dz (real)
	dcx ddzb, d2s;  // d alpha_smooth / dzbar, d^2 alpha_smooth / dzbar^2
};

// alpha_star.cuh:55-114, alpha_local.cuh:52-108, alpha_smooth.cuh:142-316. With part < parts, only
// the near stars whose position in the node's near list is part (mod parts) enter S1 and S2, and the
// expansion and smooth terms are evaluated by part 0 only (zero elsewhere): the caller adds the parts.
__device__ CCTerms cc_terms(const ShootArgs& a, int node, dcx z, bool second, int part = 0, int parts = 1)
{
	CCTerms t;
	double s1x = 0, s1y = 0, s2x = 0, s2y = 0;
	int seen = 0;
	auto range = [&](int b, int n) {
		// the first star of this range that belongs to this part
		const int first = ((part - seen) % parts + parts) % parts;
		seen += n;
		for (int s = b + first; s < b + n; s += parts)
		{
			const double dx = z.re - a.sx[s], dy = z.im - a.sy[s], m = a.sm[s];
			// 1 / (z - z_i) = conj(d) / |d|^2
			const double ir2 = dev::rcp64(dx * dx + dy * dy);
			const double ux = dx * ir2, uy = -dy * ir2;
			const double u2x = ux * ux - uy * uy, u2y = 2 * ux * uy;
			s1x += m * u2x;
			s1y += m * u2y;
			if (second)
			{
				s2x += m * (u2x * ux - u2y * uy);
				s2y += m * (u2x * uy + u2y * ux);
			}
		}
	};
	range(a.nbegin[node], a.ncount[node]);
	for (int j = 0; j < a.nnnbr[node]; j++)
	{
		const int nb = a.nnbr[8 * node + j];
		range(a.nbegin[nb], a.ncount[nb]);
	}
	t.S1 = dcx{s1x, s1y};
	t.S2 = dcx{s2x, s2y};
	t.L1 = t.L2 = t.ddzb = t.d2s = dcx{0, 0};
	t.ddz = 0;
	if (part != 0) return t;
	const double h = a.nhl[node];
	const dcx dz = dcx{(z.re - a.ncx[node]) / h, (z.im - a.ncy[node]) / h};
	const double* c = a.local + 2 * (long long)node * (a.p + 1);
	dcx l1 = {0, 0}, l2 = {0, 0};
	for (int i = a.p; i >= 2; i--)
	{
		l1 = dev::add(dev::mul(l1, dz), dev::scale(dcx{c[2 * i], c[2 * i + 1]}, double(i) * (i - 1)));
		if (second && i >= 3) l2 = dev::add(dev::mul(l2, dz), dev::scale(dcx{c[2 * i], c[2 * i + 1]}, double(i) * (i - 1) * (i - 2)));
	}
	t.L1 = dev::scale(l1, a.theta2 / (h * h));
	t.L2 = dev::scale(l2, a.theta2 / (h * h * h));
	const dcx corner = dcx{a.corner_re, a.corner_im};
	dev::d_alpha_smooth(z, a.kappastar, a.rectangular, a.approx, a.taylor_smooth, corner, t.ddz, t.ddzb);
	t.d2s = second ? dev::d2_alpha_smooth(z, a.kappastar, a.rectangular, a.approx, a.taylor_smooth, corner) : dcx{0, 0};
	return t;
}

// f(z, phi) of ccf_functions.cuh:64-80 from the terms
__device__ __forceinline__ dcx cc_f0(const ShootArgs& a, const CCTerms& t, double phi)
{
	const double dwdz = (1 - a.kappa) - t.ddz;
	double sn, cs;
	sincos(phi, &sn, &cs);
	return dcx{a.gamma + a.theta2 * t.S1.re - t.L1.re - t.ddzb.re - dwdz * cs,
		a.theta2 * t.S1.im - t.L1.im + t.ddzb.im + dwdz * sn};
}

// largest possible error in 1/mu at a root (ccf_functions.cuh:204-211, 466-472)
__device__ __forceinline__ double cc_inv_mu_error(const ShootArgs& a, const CCTerms& t, dcx f0)
{
	const double af = dev::cabs(f0), dwdz = (1 - a.kappa) - t.ddz;
	return fmax(fabs(af * (af + 2 * dwdz)), fabs(af * (af - 2 * dwdz)));
}

// row r of the root array: its branch, its step from the branch centre and its phase
struct CCRow
{
	int row;
	double phi;
};

__device__ __forceinline__ CCRow cc_row(int c, int sgn, int j, int nphi, int nb)
{
	const int J = nphi / (2 * nb);
	CCRow r;
	r.row = c * (2 * J + 1) + J + sgn * j;
	r.phi = CCF_PI / nb + c * 2 * CCF_PI / nb + sgn * j * (2 * CCF_PI / nphi);
	return r;
}

/*
One Aberth iteration of the roots at step j of every branch and side
(ccf_functions.cuh:171-240, 309-386). Each root is updated by CCF_PARTS
consecutive lanes, which share its sums over the other roots, the stars (the
poles) and its near stars, and combine them with butterfly shuffles; the
group's first lane evaluates the expansion and smooth terms and writes the
root. Block (x, y): roots x * CCF_ROOTS + [0, CCF_ROOTS) of branch y / sides,
side y % sides (side 0 steps phi down, side 1 up; at j = 0 there is one
side). src and dst hold (nphi + nb) rows of nroots roots.
*/
constexpr int CCF_PARTS = 64;
constexpr int CCF_ROOTS = CCF_THREADS / CCF_PARTS;

__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
	for (int o = CCF_PARTS / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}

__global__ __launch_bounds__(CCF_THREADS) void k_ccf_iterate(ShootArgs a, const dcx* __restrict__ src, dcx* __restrict__ dst,
	char* __restrict__ fin, int* __restrict__ leaf_cache, int nroots, int nstars, int circle_pole, int j, int nphi, int nb,
	int sides)
{
	__shared__ double tx[CCF_THREADS], ty[CCF_THREADS];
	const int c = blockIdx.y / sides, b = blockIdx.y % sides;
	const int sgn = (sides == 1) ? 0 : (b == 0 ? -1 : 1);
	const CCRow R = cc_row(c, sgn == 0 ? 1 : sgn, sgn == 0 ? 0 : j, nphi, nb);
	const dcx* row = src + (long long)R.row * nroots;
	const int part = threadIdx.x % CCF_PARTS;
	const int k = blockIdx.x * CCF_ROOTS + threadIdx.x / CCF_PARTS;
	const bool mine = k < nroots;
	const long long slot = ((long long)c * 2 + b) * nroots + k;
	char* done = fin + slot;
	const bool active = mine && !*done;
	const dcx z = mine ? row[k] : dcx{0, 0};
	// this lane's share of the sums over the other roots and over the poles (stars)
	double rx = 0, ry = 0, px = 0, py = 0;
	for (int t0 = 0; t0 < nroots; t0 += CCF_THREADS)
	{
		const int i = t0 + threadIdx.x;
		tx[threadIdx.x] = i < nroots ? row[i].re : 0;
		ty[threadIdx.x] = i < nroots ? row[i].im : 0;
		__syncthreads();
		if (active)
		{
			const int n = min(CCF_THREADS, nroots - t0);
			for (int q = part; q < n; q += CCF_PARTS)
			{
				if (t0 + q == k) continue;
				const double dx = z.re - tx[q], dy = z.im - ty[q];
				const double ir2 = dev::rcp64(dx * dx + dy * dy);
				rx += dx * ir2;
				ry -= dy * ir2;
			}
		}
		__syncthreads();
	}
	for (int t0 = 0; t0 < nstars; t0 += CCF_THREADS)
	{
		const int i = t0 + threadIdx.x;
		tx[threadIdx.x] = i < nstars ? a.sx[i] : 0;
		ty[threadIdx.x] = i < nstars ? a.sy[i] : 0;
		__syncthreads();
		if (active)
		{
			const int n = min(CCF_THREADS, nstars - t0);
			for (int q = part; q < n; q += CCF_PARTS)
			{
				const double dx = z.re - tx[q], dy = z.im - ty[q];
				const double ir2 = dev::rcp64(dx * dx + dy * dy);
				px += dx * ir2;
				py -= dy * ir2;
			}
		}
		__syncthreads();
	}
	// the whole group follows the same branch: it shares its root and its done flag
	if (!active)
	{
		if (mine && part == 0) dst[(long long)R.row * nroots + k] = z;
		return;
	}
	// the root's leaf from the previous iteration while the root stays strictly inside
	// it (there the tree descent would end in the same leaf), else a fresh descent
	const int cached = leaf_cache[slot];
	bool inside = false;
	if (cached >= 0)
	{
		const double h = a.nhl[cached] * (1 - 1e-12);
		inside = fabs(z.re - a.ncx[cached]) < h && fabs(z.im - a.ncy[cached]) < h;
	}
	const int node = inside ? cached : find_node(a, z.re, z.im);
	if (!inside && part == 0) leaf_cache[slot] = a.nchild0[node] < 0 ? node : -1;
	CCTerms t = cc_terms(a, node, z, true, part, CCF_PARTS);
	rx = group_sum(rx);
	ry = group_sum(ry);
	px = group_sum(px);
	py = group_sum(py);
	t.S1.re = group_sum(t.S1.re);
	t.S1.im = group_sum(t.S1.im);
	t.S2.re = group_sum(t.S2.re);
	t.S2.im = group_sum(t.S2.im);
	if (part != 0) return;
	dcx* out = dst + (long long)R.row * nroots + k;
	const dcx f0 = cc_f0(a, t, R.phi);
	if (cc_inv_mu_error(a, t, f0) < 1e-9)
	{
		// 1/mu is already below 1e-9 here (both bounds of ccf_functions.cuh:207-211)
		*out = z;
		*done = 1;
		return;
	}
	const dcx f1 = dcx{-2 * a.theta2 * t.S2.re - t.L2.re - t.d2s.re, -2 * a.theta2 * t.S2.im - t.L2.im + t.d2s.im};
	dcx psum = dcx{2 * px, 2 * py};
	if (circle_pole)
	{
		const dcx iz = dev::cdiv(dcx{1, 0}, z);
		psum = dev::add(psum, dev::scale(iz, 2));
	}
	const dcx den = dev::add(f1, dev::mul(f0, dcx{psum.re - rx, psum.im - ry}));
	const dcx zn = dev::sub(z, dev::cdiv(f0, den));
	*out = zn;
	if (dev::cabs(dev::sub(zn, z)) / sqrt(a.theta2) < 1e-9) *done = 1;
}

// roots of step j start at those of step j - 1 on each side; clear the done flags
__global__ void k_ccf_prepare(dcx* roots, char* fin, int nroots, int j, int nphi, int nb)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nroots) return;
	const int c = blockIdx.y / 2, b = blockIdx.y % 2;
	const int sgn = b == 0 ? -1 : 1;
	const CCRow to = cc_row(c, sgn, j, nphi, nb), from = cc_row(c, sgn, j - 1, nphi, nb);
	roots[(long long)to.row * nroots + k] = roots[(long long)from.row * nroots + k];
	fin[((long long)c * 2 + b) * nroots + k] = 0;
}

/*
Per root of every row: the error bound on 1/mu (ccf_functions.cuh:430-474),
the transposed critical curve point (root-major, as written), the caustic
point (the lens equation, 496-516) and the magnification length scale
(134-160). err_bits collects the largest error as the bit pattern of a
non-negative double; nan_flag is set by any non-finite error.
*/
__global__ void k_ccf_finish(ShootArgs a, const dcx* __restrict__ roots, int nroots, int nphi, int nb, dcx* __restrict__ ccs,
	dcx* __restrict__ caustics, double* __restrict__ mu_scales, unsigned long long* err_bits, int* nan_flag)
{
	const long long total = (long long)(nphi + nb) * nroots;
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= total) return;
	const int rowi = int(i / nroots), k = int(i % nroots);
	const int J = nphi / (2 * nb);
	const int c = rowi / (2 * J + 1), s = rowi % (2 * J + 1) - J;
	const double phi = CCF_PI / nb + c * 2 * CCF_PI / nb + s * (2 * CCF_PI / nphi);
	const dcx z = roots[i];
	const int node = find_node(a, z.re, z.im);
	const CCTerms t = cc_terms(a, node, z, mu_scales != nullptr);
	const double err = cc_inv_mu_error(a, t, cc_f0(a, t, phi));
	if (!isfinite(err))
		atomicExch(nan_flag, 1);
	else
		atomicMax(err_bits, (unsigned long long)__double_as_longlong(err));
	const long long o = (long long)k * (nphi + nb) + rowi;
	ccs[o] = z;
	if (caustics)
	{
		const Deflection d = deflect_node<false>(a, node, z.re, z.im);
		caustics[o] = dcx{d.wx, d.wy};
	}
	if (mu_scales)
	{
		// ccf_functions.cuh:134-160
		const double dwdz = (1 - a.kappa) - t.ddz;
		const dcx dwdzb = dcx{a.gamma + a.theta2 * t.S1.re - t.L1.re - t.ddzb.re, -a.theta2 * t.S1.im + t.L1.im - t.ddzb.im};
		const dcx d2 = dcx{-2 * a.theta2 * t.S2.re - t.L2.re - t.d2s.re, 2 * a.theta2 * t.S2.im + t.L2.im - t.d2s.im};
		const dcx cct = dev::mul(dev::mul(dcx{0, -2}, dev::conj(dwdzb)), d2);
		const dcx ct = dev::add(dev::scale(cct, dwdz), dev::mul(dwdzb, dev::conj(cct)));
		mu_scales[o] = 1 / (2 * dev::cabs(ct));
	}
}

}  // namespace

CCFResult Engine::ccf_solve(const LensParams& lens, const double* init_roots, int nroots, int nphi, int nb, bool circle_pole,
	double* ccs, double* caustics, double* mu_scales, int init_iters, int step_iters)
{
	if (nroots < 1 || nphi < 2 || nb < 1 || nphi % (2 * nb) != 0) throw HipError("invalid critical curve sizes", __FILE__, __LINE__);
	if (!nstars_ || !num_nodes_) throw HipError("the critical curve solver needs the tree and the FMM", __FILE__, __LINE__);
	GridParams g;
	std::memset(&g, 0, sizeof(g));
	g.num_pixels[0] = g.num_pixels[1] = 1;
	g.half_length_y[0] = g.half_length_y[1] = 1;
	ShootArgs a = ShootArgsBuilder::make(lens, g, false);
	MLENS_TREE_ARGS(a)
	const int J = nphi / (2 * nb);
	const long long nrows = nphi + nb, total = nrows * nroots;
	// complex arrays as (re, im) pairs of doubles
	DeviceBuffer<double> Ab, Bb, Cb, Wb, M;
	DeviceBuffer<char> fin;
	DeviceBuffer<int> leaf_cache;
	DeviceBuffer<unsigned long long> err;
	DeviceBuffer<int> nanf;
	Ab.alloc(2 * total);
	Bb.alloc(2 * total);
	Cb.alloc(2 * total);
	if (caustics) Wb.alloc(2 * total);
	if (mu_scales) M.alloc(total);
	struct
	{
		dcx* ptr;
	} A{reinterpret_cast<dcx*>(Ab.ptr)}, B{reinterpret_cast<dcx*>(Bb.ptr)}, C{reinterpret_cast<dcx*>(Cb.ptr)},
		W{reinterpret_cast<dcx*>(Wb.ptr)};
	fin.alloc(2 * (size_t)nb * nroots);
	leaf_cache.alloc(2 * (size_t)nb * nroots);
	err.alloc(1);
	nanf.alloc(1);
	MLENS_HIP_CHECK(hipMemsetAsync(A.ptr, 0, total * sizeof(dcx), stream_));
	MLENS_HIP_CHECK(hipMemsetAsync(fin.ptr, 0, 2 * (size_t)nb * nroots, stream_));
	MLENS_HIP_CHECK(hipMemsetAsync(leaf_cache.ptr, 0xff, 2 * (size_t)nb * nroots * sizeof(int), stream_));   // -1: none
	MLENS_HIP_CHECK(hipMemsetAsync(err.ptr, 0, sizeof(unsigned long long), stream_));
	MLENS_HIP_CHECK(hipMemsetAsync(nanf.ptr, 0, sizeof(int), stream_));
	// starting roots at the centre row of each branch
	for (int c = 0; c < nb; c++)
		MLENS_HIP_CHECK(hipMemcpyAsync(A.ptr + (long long)(c * (2 * J + 1) + J) * nroots, init_roots + 2 * (long long)c * nroots,
			nroots * sizeof(dcx), hipMemcpyHostToDevice, stream_));
	const unsigned gx = unsigned((nroots + CCF_ROOTS - 1) / CCF_ROOTS);
	// an even number of iterations leaves the roots in A
	auto iterate = [&](int j, int sides, int iters) {
		for (int it = 0; it < iters; it++)
		{
			const dcx* s = (it % 2 == 0) ? A.ptr : B.ptr;
			dcx* d = (it % 2 == 0) ? B.ptr : A.ptr;
			k_ccf_iterate<<<dim3(gx, unsigned(nb * sides)), CCF_THREADS, 0, stream_>>>(a, s, d, fin.ptr, leaf_cache.ptr,
				nroots, nstars_, circle_pole ? 1 : 0, j, nphi, nb, sides);
		}
		if (iters % 2)
		{
			// copy the rows just refined back into A
			for (int c = 0; c < nb; c++)
				for (int b = 0; b < sides; b++)
				{
					const int row = c * (2 * J + 1) + J + (sides == 1 ? 0 : (b == 0 ? -j : j));
					MLENS_HIP_CHECK(hipMemcpyAsync(A.ptr + (long long)row * nroots, B.ptr + (long long)row * nroots,
						nroots * sizeof(dcx), hipMemcpyDeviceToDevice, stream_));
				}
		}
	};
	CCFResult r;
	MLENS_HIP_CHECK(hipEventRecord(ev0_, stream_));
	iterate(0, 1, init_iters);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipEventRecord(ev1_, stream_));
	MLENS_HIP_CHECK(hipEventSynchronize(ev1_));
	float ms = 0;
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
	r.t_init_ms = ms;
	MLENS_HIP_CHECK(hipEventRecord(ev0_, stream_));
	for (int j = 1; j <= J; j++)
	{
		k_ccf_prepare<<<dim3(blocks_for(nroots, 256), unsigned(2 * nb)), 256, 0, stream_>>>(A.ptr, fin.ptr, nroots, j, nphi, nb);
		iterate(j, 2, step_iters);
	}
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipEventRecord(ev1_, stream_));
	MLENS_HIP_CHECK(hipEventSynchronize(ev1_));
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
	r.t_ccs_ms = ms;
	MLENS_HIP_CHECK(hipEventRecord(ev0_, stream_));
	k_ccf_finish<<<blocks_for(total, 256), 256, 0, stream_>>>(a, A.ptr, nroots, nphi, nb, C.ptr, caustics ? W.ptr : nullptr,
		mu_scales ? M.ptr : nullptr, err.ptr, nanf.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipEventRecord(ev1_, stream_));
	MLENS_HIP_CHECK(hipEventSynchronize(ev1_));
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
	r.t_caustics_ms = ms;
	unsigned long long eb = 0;
	int nf = 0;
	MLENS_HIP_CHECK(hipMemcpy(&eb, err.ptr, sizeof(eb), hipMemcpyDeviceToHost));
	MLENS_HIP_CHECK(hipMemcpy(&nf, nanf.ptr, sizeof(nf), hipMemcpyDeviceToHost));
	std::memcpy(&r.max_error, &eb, sizeof(double));
	r.has_nan = nf != 0;
	MLENS_HIP_CHECK(hipMemcpy(ccs, C.ptr, total * sizeof(dcx), hipMemcpyDeviceToHost));
	if (caustics) MLENS_HIP_CHECK(hipMemcpy(caustics, W.ptr, total * sizeof(dcx), hipMemcpyDeviceToHost));
	if (mu_scales) MLENS_HIP_CHECK(hipMemcpy(mu_scales, M.ptr, total * sizeof(double), hipMemcpyDeviceToHost));
	return r;
}

}  // namespace mlens
