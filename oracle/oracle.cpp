/******************************************************************************
oracle.cpp -- TEST INFRASTRUCTURE ONLY. NOT PART OF THE PRODUCT.

A CPU, FP64, OpenMP restatement of the weisluke/irs IPM/IRS map-making
algorithm. It is the checker for the HIP engine in irs_amd/ and the reported
CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may load it, and only as the checker. The product never links or calls it.

Every function cites the reference file:line it restates. The reference is
CUDA-only and cannot be built here (no nvcc), so the restatement is pinned by
closed forms, brute force and published known-answer vectors in tests/.

Deliberate differences from the reference:
  * binomial coefficients are exact (uint64, rounded once to double) instead
    of int32, which wraps for rows >= 34 (binomial_coefficients.cuh:11-44);
  * star fields come from Philox4x32-10 (Salmon et al., SC'11) instead of
    curand XORWOW, whose internals are not reproduced;
  * a cell whose corner deflects to NaN drops exactly the triangles that use
    that corner (the reference's `break` at ipm_functions.cuh:153 leaves the
    later corners at (0,0));
  * child order inside a node and star order inside a child are deterministic
    (stable partition) instead of atomic-order dependent.
******************************************************************************/

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

/* ---------------------------------------------------------------------------
Complex arithmetic, restating complex.cuh. Division uses Smith's formula
(R. L. Smith, CACM 5(8), 1962), as complex.cuh:247-280 does.
--------------------------------------------------------------------------- */
struct Cx
{
	double re;
	double im;
};

inline Cx cx(double r, double i = 0) { return Cx{r, i}; }
inline Cx operator+(Cx a, Cx b) { return Cx{a.re + b.re, a.im + b.im}; }
inline Cx operator-(Cx a, Cx b) { return Cx{a.re - b.re, a.im - b.im}; }
inline Cx operator-(Cx a) { return Cx{-a.re, -a.im}; }
inline Cx operator*(Cx a, Cx b) { return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx operator*(Cx a, double s) { return Cx{a.re * s, a.im * s}; }
inline Cx operator*(double s, Cx a) { return Cx{s * a.re, s * a.im}; }
inline Cx operator/(Cx a, double s) { return Cx{a.re / s, a.im / s}; }
inline Cx conj(Cx a) { return Cx{a.re, -a.im}; }
inline double cabs(Cx a) { return std::hypot(a.re, a.im); }           // complex.cuh:81-84
inline double carg(Cx a) { return std::atan2(a.im, a.re); }           // complex.cuh:89-92
inline Cx clog(Cx a) { return Cx{std::log(cabs(a)), carg(a)}; }       // complex.cuh:105-111
inline Cx cexp(Cx a) { return Cx{std::exp(a.re) * std::cos(a.im), std::exp(a.re) * std::sin(a.im)}; }

// complex.cuh:247-262
inline Cx operator/(Cx a, Cx b)
{
	if (std::abs(b.im) < std::abs(b.re))
	{
		double f = b.im / b.re;
		return Cx{(a.re + a.im * f) / (b.re + b.im * f), (a.im - a.re * f) / (b.re + b.im * f)};
	}
	double f = b.re / b.im;
	return Cx{(a.re * f + a.im) / (b.re * f + b.im), (a.im * f - a.re) / (b.re * f + b.im)};
}

// complex.cuh:267-280
inline Cx operator/(double a, Cx b)
{
	if (std::abs(b.im) < std::abs(b.re))
	{
		double f = b.im / b.re;
		return Cx{a / (b.re + b.im * f), (-a * f) / (b.re + b.im * f)};
	}
	double f = b.re / b.im;
	return Cx{(a * f) / (b.re * f + b.im), -a / (b.re * f + b.im)};
}

// complex.cuh:285-306: integer power by repeated multiplication
inline Cx cpow(Cx a, int n)
{
	Cx r = cx(1, 0);
	if (n > 0)
	{
		for (int i = 0; i < n; i++) r = r * a;
	}
	else if (n < 0)
	{
		for (int i = 0; i > n; i--) r = r * a;
		r = cx(1, 0) / r;
	}
	return r;
}

constexpr double PI = 3.14159265358979323846;

/* ---------------------------------------------------------------------------
Mass functions: mass_functions/power_law.cuh:30-226, kroupa.cuh:22-502,
salpeter.cuh:22, uniform.cuh, equal.cuh:27-72, optical_depth.cuh:28-91.
--------------------------------------------------------------------------- */
enum MassFunctionId { MF_EQUAL = 0, MF_UNIFORM = 1, MF_SALPETER = 2, MF_KROUPA = 3, MF_OPTICAL_DEPTH = 4 };

template <typename T>
T power_integral(T x1, T x2, T a, T b = 1)                                      // power_law.cuh:30-44
{
	if (a != -1) return b * (std::pow(x2, a + 1) - std::pow(x1, a + 1)) / (a + 1);
	return b * (std::log(x2) - std::log(x1));
}
template <typename T>
T power_log_integral(T x1, T x2, T a, T b = 1)                                  // power_law.cuh:59-73
{
	if (a != -1)
		return b * (std::pow(x2, a + 1) * ((a + 1) * std::log(x2) - 1) - std::pow(x1, a + 1) * ((a + 1) * std::log(x1) - 1)) / ((a + 1) * (a + 1));
	return b * (std::log(x2) * std::log(x2) - std::log(x1) * std::log(x1)) / 2;
}
double invert_power_integral(double p, double x1, double a, double b)           // power_law.cuh:87-101
{
	if (a != -1) return std::pow(p * (a + 1) / b + std::pow(x1, a + 1), 1 / (a + 1));
	return x1 * std::exp(p / b);
}

/*
A broken power law restricted to [ml, mu]: the segments of the Kroupa
function (kroupa.cuh:22-33) that intersect the mass range, with the pdf
continuous across the breaks. The reference enumerates the same segment
combinations branch by branch (kroupa.cuh:63-123); a segment is present iff
it has positive width, which is exactly the reference's strict comparisons.
*/
template <typename T>
struct SegmentsT
{
	int n = 0;
	T lo[3], hi[3], slope[3], b[3], prob[3];
};
using Segments = SegmentsT<double>;

template <typename T>
SegmentsT<T> make_segments(int mf, T ml, T mu, T ms)
{
	SegmentsT<T> s;
	if (mf == MF_KROUPA)
	{
		const T slopes[3] = {T(-0.3), T(-1.3), T(-2.3)};
		const T m0 = T(0.08) * ms;
		const T m1 = T(0.5) * ms;
		const T edges[4] = {0, m0, m1, std::numeric_limits<T>::infinity()};
		T k = 1;        // continuity factor relative to the first present segment
		bool first = true;
		T total = 0;
		for (int i = 0; i < 3; i++)
		{
			T lo = std::max(ml, edges[i]);
			T hi = std::min(mu, edges[i + 1]);
			if (!(lo < hi)) continue;
			if (!first) k *= std::pow(edges[i], slopes[i - 1] - slopes[i]);
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
		T slope = (mf == MF_SALPETER) ? T(-2.35) : T(0);
		s.n = 1;
		s.lo[0] = ml;
		s.hi[0] = mu;
		s.slope[0] = slope;
		s.b[0] = 1 / power_integral<T>(ml, mu, slope);
		s.prob[0] = 1;
	}
	return s;
}

double mass_sample(int mf, double p, double ml, double mu, double ms)
{
	switch (mf)
	{
	case MF_EQUAL:
		return 1;                                                                 // equal.cuh:27-30
	case MF_OPTICAL_DEPTH:                                                        // optical_depth.cuh:28-40
		if (ml == mu) return ml;
		return (p <= mu / (ml + mu)) ? ml : mu;
	default:
	{
		if (ml == mu) return ml;                                                  // power_law.cuh:128-131
		Segments s = make_segments<double>(mf, ml, mu, ms);
		double cum = 0;
		for (int i = 0; i < s.n; i++)
		{
			if (p < cum + s.prob[i] || i == s.n - 1)
			{
				return invert_power_integral(p - cum, s.lo[i], s.slope[i], s.b[i]);  // kroupa.cuh:130-159
			}
			cum += s.prob[i];
		}
		return s.lo[0];
	}
	}
}

/* the moments <m>, <m^2>, <m^2 ln m> in the value type T, as the reference's
   mass functions evaluate them (mass_function<T>) */
template <typename T>
void mass_moments(int mf, T ml, T mu, T ms, T out[3])
{
	switch (mf)
	{
	case MF_EQUAL:                                                                // equal.cuh:41-72
		out[0] = 1;
		out[1] = 1;
		out[2] = 0;
		return;
	case MF_OPTICAL_DEPTH:                                                        // optical_depth.cuh:49-91
		if (ml == mu)
		{
			out[0] = ml;
			out[1] = ml * ml;
			out[2] = ml * ml * std::log(ml);
			return;
		}
		out[0] = 2 * ml * mu / (ml + mu);
		out[1] = ml * mu;
		out[2] = ml * mu * (ml * std::log(ml) + mu * std::log(mu)) / (ml + mu);
		return;
	default:
	{
		if (ml == mu)
		{
			out[0] = ml;
			out[1] = ml * ml;
			out[2] = ml * ml * std::log(ml);
			return;
		}
		SegmentsT<T> s = make_segments<T>(mf, ml, mu, ms);
		out[0] = out[1] = out[2] = 0;
		for (int i = 0; i < s.n; i++)
		{
			out[0] += power_integral<T>(s.lo[i], s.hi[i], s.slope[i] + 1, s.b[i]);
			out[1] += power_integral<T>(s.lo[i], s.hi[i], s.slope[i] + 2, s.b[i]);
			out[2] += power_log_integral<T>(s.lo[i], s.hi[i], s.slope[i] + 2, s.b[i]);
		}
	}
	}
}

/* ---------------------------------------------------------------------------
Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11). Published multipliers and Weyl
constants.
--------------------------------------------------------------------------- */
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
	uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
	uint32_t k0 = key_in[0];
	uint32_t k1 = key_in[1];
	for (int r = 0; r < 10; r++)
	{
		if (r > 0)
		{
			k0 += 0x9E3779B9u;
			k1 += 0xBB67AE85u;
		}
		uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
		uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
		uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
		uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
		c[0] = hi1 ^ c[1] ^ k0;
		c[1] = lo1;
		c[2] = hi0 ^ c[3] ^ k1;
		c[3] = lo0;
	}
	for (int i = 0; i < 4; i++) out[i] = c[i];
}

// 53 random bits from two words, mapped to (0, 1]
inline double u01(uint32_t a, uint32_t b)
{
	uint64_t x = ((uint64_t(a) << 32) | b) >> 11;
	return double(x + 1) * (1.0 / 9007199254740992.0);
}

/* star i draws uniforms from counters (i, 0) and (i, 1) under key (seed, 0) */
inline void star_uniforms(uint32_t seed, int64_t i, double u[4])
{
	uint32_t key[2] = {seed, 0u};
	uint32_t ctr[4] = {uint32_t(i), uint32_t(uint64_t(i) >> 32), 0u, 0u};
	uint32_t o[4];
	philox4x32_10(ctr, key, o);
	u[0] = u01(o[0], o[1]);
	u[1] = u01(o[2], o[3]);
	ctr[2] = 1u;
	philox4x32_10(ctr, key, o);
	u[2] = u01(o[0], o[1]);
	u[3] = u01(o[2], o[3]);
}

/* ---------------------------------------------------------------------------
Quadtree: tree_node.cuh:32-561 and the level loop of ipm.cuh:740-839.
--------------------------------------------------------------------------- */
constexpr int MAX_NUM_STARS_DIRECT = 32;   // tree_node.cuh:14
constexpr int MAX_EXPANSION_ORDER = 31;    // tree_node.cuh:20

struct Node
{
	double cx, cy, hl;
	int level;
	int parent;
	int child0;      // index of child 0; children are consecutive; -1 if none
	int nchild;
	int begin, n;    // star range
	int nnbr, nm2l, np2l;
	int nbr[8];
	int m2l[27];
	int p2l[5];
};

struct Tree
{
	std::vector<double> sx, sy, sm;   // stars, sorted into tree order
	std::vector<Node> nodes;
	std::vector<int> level_start;     // nodes of level l are [level_start[l], level_start[l+1])
	int levels = 0;                   // tree_levels of the reference
	int p = 0;
	std::vector<Cx> mult;             // (p+1) per node
	std::vector<Cx> local;            // (p+1) per node
	std::vector<double> binom;        // Pascal triangle rows 0..2p, triangular indexing
};

inline double binom(const Tree& t, int n, int k)           // binomial_coefficients.cuh:53-64
{
	if (n < k) return 0;
	return t.binom[size_t(n) * (n + 1) / 2 + k];
}

// tree_node.cuh:183-267
void set_neighbors(Tree& t, int idx)
{
	Node& me = t.nodes[idx];
	if (me.level == 0) return;
	const Node& par = t.nodes[me.parent];
	for (int i = 0; i < par.nchild; i++)
	{
		int c = par.child0 + i;
		if (t.nodes[c].n == 0) continue;
		if (c != idx) me.nbr[me.nnbr++] = c;
	}
	for (int i = 0; i < par.nnbr; i++)
	{
		int nb = par.nbr[i];
		const Node& N = t.nodes[nb];
		if (N.n == 0) continue;
		if (N.nchild == 0)
		{
			if (std::fabs(N.cx - me.cx) < 3.5 * me.hl && std::fabs(N.cy - me.cy) < 3.5 * me.hl)
				me.nbr[me.nnbr++] = nb;
			else
				me.p2l[me.np2l++] = nb;
		}
		else
		{
			for (int j = 0; j < N.nchild; j++)
			{
				int c = N.child0 + j;
				const Node& C = t.nodes[c];
				if (C.n == 0) continue;
				if (std::fabs(C.cx - me.cx) < 3 * me.hl && std::fabs(C.cy - me.cy) < 3 * me.hl)
					me.nbr[me.nnbr++] = c;
				else
					me.m2l[me.nm2l++] = c;
			}
		}
	}
}

inline int quadrant(double x, double y, double cx, double cy)     // tree_node.cuh:432-477, 542-557
{
	if (y >= cy) return (x >= cx) ? 0 : 1;
	return (x < cx) ? 2 : 3;
}

void build_tree(Tree& t, double root_hl)
{
	t.nodes.clear();
	t.level_start.clear();
	Node root{};
	root.cx = 0;
	root.cy = 0;
	root.hl = root_hl;
	root.level = 0;
	root.parent = -1;
	root.child0 = -1;
	root.begin = 0;
	root.n = int(t.sx.size());
	t.nodes.push_back(root);
	t.level_start.push_back(0);
	t.level_start.push_back(1);
	t.levels = 0;

	std::vector<double> tx(t.sx.size()), ty(t.sx.size()), tm(t.sx.size());
	while (true)
	{
		int l0 = t.level_start[t.levels];
		int l1 = t.level_start[t.levels + 1];
		// ipm.cuh:797-806 with tree_node.cuh:321-345
		int maxn = 0;
		std::vector<char> split(l1 - l0, 0);
		for (int i = l0; i < l1; i++)
		{
			const Node& N = t.nodes[i];
			int cnt = N.n;
			for (int j = 0; j < N.nnbr; j++) cnt += t.nodes[N.nbr[j]].n;
			maxn = std::max(maxn, cnt);
			split[i - l0] = cnt > MAX_NUM_STARS_DIRECT;
		}
		if (maxn <= MAX_NUM_STARS_DIRECT) break;
		// tree_node.cuh:151-177, 355-381, 392-501
		for (int i = l0; i < l1; i++)
		{
			if (!split[i - l0]) continue;
			int c0 = int(t.nodes.size());
			Node par = t.nodes[i];
			double nh = par.hl / 2;
			const double off[4][2] = {{1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
			int cnt[4] = {0, 0, 0, 0};
			for (int s = par.begin; s < par.begin + par.n; s++) cnt[quadrant(t.sx[s], t.sy[s], par.cx, par.cy)]++;
			int start[4];
			start[0] = par.begin;
			for (int q = 1; q < 4; q++) start[q] = start[q - 1] + cnt[q - 1];
			int pos[4] = {start[0], start[1], start[2], start[3]};
			for (int s = par.begin; s < par.begin + par.n; s++)
			{
				int q = quadrant(t.sx[s], t.sy[s], par.cx, par.cy);
				tx[pos[q]] = t.sx[s];
				ty[pos[q]] = t.sy[s];
				tm[pos[q]] = t.sm[s];
				pos[q]++;
			}
			for (int s = par.begin; s < par.begin + par.n; s++)
			{
				t.sx[s] = tx[s];
				t.sy[s] = ty[s];
				t.sm[s] = tm[s];
			}
			for (int q = 0; q < 4; q++)
			{
				Node c{};
				c.cx = par.cx + nh * off[q][0];
				c.cy = par.cy + nh * off[q][1];
				c.hl = nh;
				c.level = par.level + 1;
				c.parent = i;
				c.child0 = -1;
				c.begin = start[q];
				c.n = cnt[q];
				t.nodes.push_back(c);
			}
			t.nodes[i].child0 = c0;
			t.nodes[i].nchild = 4;
		}
		t.levels++;
		t.level_start.push_back(int(t.nodes.size()));
		int n0 = t.level_start[t.levels];
		int n1 = t.level_start[t.levels + 1];
#pragma omp parallel for schedule(static)
		for (int i = n0; i < n1; i++) set_neighbors(t, i);
	}
}

// tree_node.cuh:529-561
int nearest_node(const Tree& t, double x, double y)
{
	int idx = 0;
	while (t.nodes[idx].nchild > 0)
	{
		const Node& N = t.nodes[idx];
		if (std::fabs(x - N.cx) > N.hl || std::fabs(y - N.cy) > N.hl) return idx;
		idx = N.child0 + quadrant(x, y, N.cx, N.cy);
	}
	return idx;
}

/* ---------------------------------------------------------------------------
FMM in node-normalised units: fmm.cuh:21-464.
--------------------------------------------------------------------------- */
void compute_binomials(Tree& t, int nmax)
{
	// exact integers in uint64 (C(62,31) < 2^64), rounded once to double
	std::vector<uint64_t> row(nmax + 1, 0), prev(nmax + 1, 0);
	t.binom.assign(size_t(nmax + 1) * (nmax + 2) / 2, 0.0);
	for (int n = 0; n <= nmax; n++)
	{
		for (int k = 0; k <= n; k++) row[k] = (k == 0 || k == n) ? 1 : prev[k - 1] + prev[k];
		for (int k = 0; k <= n; k++) t.binom[size_t(n) * (n + 1) / 2 + k] = double(row[k]);
		prev = row;
	}
}

void fmm(Tree& t, int p)
{
	t.p = p;
	compute_binomials(t, 2 * p);
	size_t nn = t.nodes.size();
	t.mult.assign(nn * (p + 1), cx(0));
	t.local.assign(nn * (p + 1), cx(0));

	for (int lev = t.levels; lev >= 0; lev--)
	{
		int l0 = t.level_start[lev];
		int l1 = t.level_start[lev + 1];
		// P2M, fmm.cuh:21-98 (leaves that hold stars)
#pragma omp parallel for schedule(dynamic, 64)
		for (int i = l0; i < l1; i++)
		{
			const Node& N = t.nodes[i];
			if (N.nchild != 0 || N.n == 0) continue;
			Cx* a = &t.mult[size_t(i) * (p + 1)];
			for (int k = 0; k <= p; k++)
			{
				Cx r = cx(0);
				if (k == 0)
				{
					for (int s = N.begin; s < N.begin + N.n; s++) r = r + cx(t.sm[s]);
				}
				else
				{
					for (int s = N.begin; s < N.begin + N.n; s++)
					{
						Cx dz = cx(t.sx[s] - N.cx, t.sy[s] - N.cy) / N.hl;
						r = r - cpow(dz, k) * t.sm[s];
					}
					r = r / double(k);
				}
				a[k] = r;
			}
		}
		// M2M, fmm.cuh:109-192 (children to parent)
#pragma omp parallel for schedule(dynamic, 64)
		for (int i = l0; i < l1; i++)
		{
			const Node& N = t.nodes[i];
			if (N.nchild == 0) continue;
			Cx* b = &t.mult[size_t(i) * (p + 1)];
			for (int c = 0; c < N.nchild; c++)
			{
				int ci = N.child0 + c;
				const Node& C = t.nodes[ci];
				const Cx* a = &t.mult[size_t(ci) * (p + 1)];
				Cx dz = cx(C.cx - N.cx, C.cy - N.cy) / C.hl;
				for (int l = 0; l <= p; l++)
				{
					Cx r = cx(0);
					if (l == 0)
					{
						r = a[0];
					}
					else
					{
						for (int k = l; k >= 1; k--)
						{
							r = r + a[k] * binom(t, l - 1, k - 1);
							r = r / dz;
						}
						r = r - a[0] / double(l);
						r = r * cpow(dz, l);
						r = r / std::ldexp(1.0, l);
					}
					b[l] = b[l] + r;
				}
			}
		}
	}

	for (int lev = 2; lev <= t.levels; lev++)
	{
		int l0 = t.level_start[lev];
		int l1 = t.level_start[lev + 1];
#pragma omp parallel for schedule(dynamic, 64)
		for (int i = l0; i < l1; i++)
		{
			const Node& N = t.nodes[i];
			Cx* d = &t.local[size_t(i) * (p + 1)];
			// L2L, fmm.cuh:204-268 (overwrites)
			{
				const Node& P = t.nodes[N.parent];
				const Cx* c = &t.local[size_t(N.parent) * (p + 1)];
				Cx dz = cx(N.cx - P.cx, N.cy - P.cy) / P.hl;
				for (int l = 0; l <= p; l++)
				{
					Cx r = cx(0);
					for (int k = p; k >= 1; k--)
					{
						r = r + c[k] * binom(t, k, l);
						r = r * dz;
					}
					r = r / cpow(dz, l);
					r = r / std::ldexp(1.0, l);
					d[l] = r;
				}
			}
			// M2L, fmm.cuh:281-374
			for (int j = 0; j < N.nm2l; j++)
			{
				const Node& F = t.nodes[N.m2l[j]];
				const Cx* b = &t.mult[size_t(N.m2l[j]) * (p + 1)];
				Cx dz = cx(F.cx - N.cx, F.cy - N.cy) / F.hl;
				Cx mdz = -dz;
				for (int l = 0; l <= p; l++)
				{
					Cx r = cx(0);
					if (l == 0)
					{
						for (int k = p; k >= 1; k--)
						{
							r = r + b[k];
							r = r / mdz;
						}
						r = r + b[0] * clog(mdz);
						r = r + b[0] * std::log(F.hl);
					}
					else
					{
						for (int k = p; k >= 1; k--)
						{
							r = r + b[k] * binom(t, l + k - 1, k - 1);
							r = r / mdz;
						}
						r = r - b[0] / double(l);
						r = r / cpow(dz, l);
					}
					d[l] = d[l] + r;
				}
			}
			// P2L, fmm.cuh:387-464
			for (int j = 0; j < N.np2l; j++)
			{
				const Node& F = t.nodes[N.p2l[j]];
				for (int l = 0; l <= p; l++)
				{
					Cx r = cx(0);
					for (int s = F.begin; s < F.begin + F.n; s++)
					{
						Cx dz = cx(t.sx[s] - N.cx, t.sy[s] - N.cy) / N.hl;
						if (l == 0)
						{
							r = r + t.sm[s] * clog(-dz);
							r = r + cx(t.sm[s] * std::log(N.hl));
						}
						else
						{
							r = r - cx(t.sm[s]) / (double(l) * cpow(dz, l));
						}
					}
					d[l] = d[l] + r;
				}
			}
		}
	}
}

/* ---------------------------------------------------------------------------
Deflection: alpha_star.cuh:18-77, alpha_local.cuh:18-69,
alpha_smooth.cuh:15-236, ipm_functions.cuh:34-82.
--------------------------------------------------------------------------- */
struct Lens
{
	double kappa, gamma, theta, kappastar;
	int rectangular;
	double corner_re, corner_im;
	int approx;
	int taylor_smooth;
};

Cx alpha_smooth(Cx z, const Lens& L)                                           // alpha_smooth.cuh:62-127
{
	Cx corner = cx(L.corner_re, L.corner_im);
	double ks = L.kappastar;
	Cx a;
	if (L.rectangular)
	{
		if (L.approx)
		{
			Cx s1 = cx(0), s2 = cx(0), s3 = cx(0), s4 = cx(0);
			for (int i = L.taylor_smooth; i >= 1; i--)
			{
				s1 = s1 + cx(1.0 / i);
				s2 = s2 + cx(1.0 / i);
				s3 = s3 + cx(1.0 / i);
				s4 = s4 + cx(1.0 / i);
				s1 = s1 * (conj(z) / corner);
				s2 = s2 * (conj(z) / conj(corner));
				s3 = s3 * (conj(z) / -corner);
				s4 = s4 * (conj(z) / -conj(corner));
			}
			a = (corner - conj(z)) * (clog(corner) - s1) - (conj(corner) - conj(z)) * (clog(conj(corner)) - s2)
				+ (-corner - conj(z)) * (clog(-corner) - s3) - (-conj(corner) - conj(z)) * (clog(conj(-corner)) - s4);
			a = a * cx(0, -ks / PI);
			a = a - cx(ks * 2 * (corner.re + z.re));
		}
		else
		{
			Cx c1 = corner - conj(z);
			Cx c2 = conj(corner) - conj(z);
			Cx c3 = -corner - conj(z);
			Cx c4 = -conj(corner) - conj(z);
			a = c1 * clog(c1) - c2 * clog(c2) + c3 * clog(c3) - c4 * clog(c4);
			a = a * cx(0, -ks / PI);
			double box = (-corner.re <= z.re && z.re <= corner.re && -corner.im <= z.im && z.im <= corner.im) ? 1 : 0;
			a = a - cx(ks * 2 * (corner.re + z.re) * box);
			double h1 = (corner.im + z.im >= 0) ? 1 : 0;
			double h2 = (corner.im - z.im >= 0) ? 1 : 0;
			double h3 = (z.re - corner.re >= 0) ? 1 : 0;
			a = a - cx(ks * 4 * corner.re * h1 * h2 * h3);
		}
	}
	else
	{
		if (L.approx)
		{
			a = -ks * z;
		}
		else
		{
			if (cabs(z) <= cabs(corner))
				a = -ks * z;
			else
				a = -ks * cabs(corner) * cabs(corner) / conj(z);
		}
	}
	return a;
}

double d_alpha_smooth_d_z(Cx z, const Lens& L)                                // alpha_smooth.cuh:141-170
{
	Cx corner = cx(L.corner_re, L.corner_im);
	if (L.rectangular)
	{
		if (L.approx) return -L.kappastar;
		double box = (-corner.re <= z.re && z.re <= corner.re && -corner.im <= z.im && z.im <= corner.im) ? 1 : 0;
		return -L.kappastar * box;
	}
	if (L.approx) return -L.kappastar;
	return -L.kappastar * ((cabs(corner) - cabs(z) >= 0) ? 1 : 0);
}

Cx d_alpha_smooth_d_zbar(Cx z, const Lens& L)                                 // alpha_smooth.cuh:185-236
{
	Cx corner = cx(L.corner_re, L.corner_im);
	double ks = L.kappastar;
	Cx d = cx(0);
	if (L.rectangular)
	{
		if (L.approx)
		{
			Cx r = conj(z) / corner;
			Cx phase = cx(0, carg(corner));
			for (int i = L.taylor_smooth - 1; i >= 2; i -= 2)
			{
				d = d + (cx(1) - cexp(2.0 * phase * double(i))) / double(i);
				d = d * (r * r);
			}
			d = d * 2.0;
			d = d * cx(0, -ks / PI);
			d = d + cx(ks - 4 * ks * carg(corner) / PI);
		}
		else
		{
			Cx c1 = conj(corner) - conj(z);
			Cx c2 = corner - conj(z);
			Cx c3 = -corner - conj(z);
			Cx c4 = -conj(corner) - conj(z);
			d = clog(c1) - clog(c2) - clog(c3) + clog(c4);
			d = d * cx(0, -ks / PI);
			double box = (-corner.re <= z.re && z.re <= corner.re && -corner.im <= z.im && z.im <= corner.im) ? 1 : 0;
			d = d - cx(ks * box);
		}
	}
	else
	{
		if (!L.approx)
		{
			Cx z2 = cpow(conj(z), 2);
			d = cx(ks * cabs(corner) * cabs(corner)) / z2;
		}
	}
	return d;
}

struct DeflectOut
{
	Cx w;
	double mu;
};

// ipm_functions.cuh:34-46 (and 67-82 when want_mu)
DeflectOut deflect_tree(const Tree& t, const Lens& L, Cx z, bool want_mu, int* leaf_out = nullptr)
{
	int idx = nearest_node(t, z.re, z.im);
	if (leaf_out) *leaf_out = idx;
	const Node& N = t.nodes[idx];
	// alpha_star.cuh:18-41
	Cx as = cx(0);
	Cx das = cx(0);
	auto add_range = [&](int b, int n) {
		for (int s = b; s < b + n; s++)
		{
			Cx d = z - cx(t.sx[s], t.sy[s]);
			as = as + cx(t.sm[s]) / d;
			if (want_mu) das = das + cx(t.sm[s]) / (d * d);
		}
	};
	for (int j = 0; j < N.nnbr; j++) add_range(t.nodes[N.nbr[j]].begin, t.nodes[N.nbr[j]].n);
	add_range(N.begin, N.n);
	double th2 = L.theta * L.theta;
	as = conj(as * th2);
	// alpha_local.cuh:18-36; the node's expansion order is p whenever any
	// coefficient was written (tree_node.cuh:269-303), else the loop is empty
	Cx al = cx(0);
	Cx dal = cx(0);
	if (t.p > 0)
	{
		const Cx* c = &t.local[size_t(idx) * (t.p + 1)];
		Cx dz = cx(z.re - N.cx, z.im - N.cy) / N.hl;
		for (int i = t.p; i >= 1; i--)
		{
			al = al * dz;
			al = al + c[i] * double(i);
		}
		al = conj(al * th2 / N.hl);
		if (want_mu)
		{
			for (int i = t.p; i >= 2; i--)
			{
				dal = dal * dz;
				dal = dal + c[i] * double(i * (i - 1));
			}
			dal = conj(dal * th2 / (N.hl * N.hl));
		}
	}
	Cx asm_ = alpha_smooth(z, L);
	DeflectOut o;
	o.w = (1 - L.kappa) * z + L.gamma * conj(z) - as - al - asm_;
	o.mu = 0;
	if (want_mu)
	{
		das = conj(das * (-th2));
		double dwdz = (1 - L.kappa) - d_alpha_smooth_d_z(z, L);
		Cx dwdzb = cx(L.gamma) - das - dal - d_alpha_smooth_d_zbar(z, L);
		double muinv = dwdz * dwdz - cabs(dwdzb) * cabs(dwdzb);
		o.mu = 1 / muinv;
	}
	return o;
}

/* ---------------------------------------------------------------------------
Triangle allocation among pixels: polygon.cuh:32-467, with the reference's
log/exp form of the edge intersection (polygon.cuh:150-197).
--------------------------------------------------------------------------- */
constexpr int MAX_NUM_SIDES = 7;   // polygon.cuh:11

struct Poly
{
	Cx pt[MAX_NUM_SIDES];
	int n = 0;
	void add(Cx p)                                                             // polygon.cuh:95-111
	{
		if (n >= MAX_NUM_SIDES)
		{
			n = 0;
			return;
		}
		pt[n++] = p;
	}
	double area() const                                                         // polygon.cuh:116-144
	{
		if (n < 3) return 0;
		double r = 0;
		for (int i = 0; i < n; i++)
		{
			double dx = pt[(i + 1) % n].re - pt[i].re;
			double dy = pt[(i + 1) % n].im - pt[i].im;
			r += dy * pt[i].re - dx * pt[i].im;
		}
		return r / 2;
	}
	double minx() const { double m = pt[0].re; for (int i = 1; i < n; i++) m = std::min(m, pt[i].re); return m; }
	double maxx() const { double m = pt[0].re; for (int i = 1; i < n; i++) m = std::max(m, pt[i].re); return m; }
	double miny() const { double m = pt[0].im; for (int i = 1; i < n; i++) m = std::min(m, pt[i].im); return m; }
	double maxy() const { double m = pt[0].im; for (int i = 1; i < n; i++) m = std::max(m, pt[i].im); return m; }
};

inline double sgn(double v) { return (v < -0.0) ? -1 : ((v > 0) ? 1 : 0); }    // polygon.cuh:21-27

double x_intersection(double y, Cx p1, Cx p2)                                   // polygon.cuh:150-173
{
	double dx = p2.re - p1.re;
	if (dx == 0) return p1.re;
	double ldx = std::log(std::fabs(dx));
	double dy = p2.im - p1.im;
	double ldy = std::log(std::fabs(dy));
	double lt = std::log(std::fabs(y - p1.im)) - ldy;
	return p1.re + sgn(y - p1.im) * sgn(dy) * sgn(dx) * std::exp(lt + ldx);
}
double y_intersection(double x, Cx p1, Cx p2)                                   // polygon.cuh:174-197
{
	double dy = p2.im - p1.im;
	if (dy == 0) return p1.im;
	double ldy = std::log(std::fabs(dy));
	double dx = p2.re - p1.re;
	double ldx = std::log(std::fabs(dx));
	double lt = std::log(std::fabs(x - p1.re)) - ldx;
	return p1.im + sgn(x - p1.re) * sgn(dx) * sgn(dy) * std::exp(lt + ldy);
}

// polygon.cuh:254-300: returns the left part, keeps the right (clip_left) or vice versa
Poly clip_at_x(Poly& P, double x, bool clip_left)
{
	Poly L, R;
	for (int i = 0; i < P.n; i++)
	{
		Cx a = P.pt[i];
		Cx b = P.pt[(i + 1) % P.n];
		if (a.re < x)
		{
			L.add(a);
			if (b.re > x)
			{
				double y = y_intersection(x, a, b);
				L.add(cx(x, y));
				R.add(cx(x, y));
			}
		}
		else if (a.re > x)
		{
			R.add(a);
			if (b.re < x)
			{
				double y = y_intersection(x, a, b);
				L.add(cx(x, y));
				R.add(cx(x, y));
			}
		}
		else
		{
			L.add(a);
			R.add(a);
		}
	}
	if (clip_left)
	{
		P = R;
		return L;
	}
	P = L;
	return R;
}
// polygon.cuh:301-347
Poly clip_at_y(Poly& P, double y, bool clip_bottom)
{
	Poly B, T;
	for (int i = 0; i < P.n; i++)
	{
		Cx a = P.pt[i];
		Cx b = P.pt[(i + 1) % P.n];
		if (a.im < y)
		{
			B.add(a);
			if (b.im > y)
			{
				double x = x_intersection(y, a, b);
				B.add(cx(x, y));
				T.add(cx(x, y));
			}
		}
		else if (a.im > y)
		{
			T.add(a);
			if (b.im < y)
			{
				double x = x_intersection(y, a, b);
				B.add(cx(x, y));
				T.add(cx(x, y));
			}
		}
		else
		{
			B.add(a);
			T.add(a);
		}
	}
	if (clip_bottom)
	{
		P = T;
		return B;
	}
	P = B;
	return T;
}

// polygon.cuh:349-467; emit(index, value) receives each pixel contribution
template <class Emit>
bool allocate_area_among_pixels(Poly P, double factor, int n1, int n2, Emit emit)
{
	if (P.n < 3) return false;
	if (P.minx() >= n1 || P.maxx() <= 0 || P.miny() >= n2 || P.maxy() <= 0) return false;
	double whole = std::fabs(P.area());
	if (whole == 0) return false;
	if (P.minx() < 0) clip_at_x(P, 0, true);
	if (P.maxx() > n1) clip_at_x(P, n1, false);
	if (P.miny() < 0) clip_at_y(P, 0, true);
	if (P.maxy() > n2) clip_at_y(P, n2, false);
	if (P.n < 3) return false;
	int xmin = int(P.minx());
	int xmax = int(P.maxx());
	if (xmax == n1) xmax--;
	for (int x = xmin; x < xmax; x++)
	{
		Poly top = clip_at_x(P, x + 1, true);
		if (top.n < 3) continue;
		int ymin = int(top.miny());
		int ymax = int(top.maxy());
		if (ymax == n2) ymax--;
		for (int y = ymin; y < ymax; y++)
		{
			Poly bot = clip_at_y(top, y + 1, true);
			if (bot.n < 3) continue;
			emit(size_t(n1) * y + x, factor * std::fabs(bot.area()) / whole);
		}
		emit(size_t(n1) * ymax + x, factor * std::fabs(top.area()) / whole);
	}
	if (P.n < 3) return true;
	int ymin = int(P.miny());
	int ymax = int(P.maxy());
	if (ymax == n2) ymax--;
	for (int y = ymin; y < ymax; y++)
	{
		Poly bot = clip_at_y(P, y + 1, true);
		if (bot.n < 3) continue;
		emit(size_t(n1) * y + xmax, factor * std::fabs(bot.area()) / whole);
	}
	emit(size_t(n1) * ymax + xmax, factor * std::fabs(P.area()) / whole);
	return true;
}

inline void atomic_add(double* p, double v)
{
#pragma omp atomic
	*p += v;
}

}  // namespace

/* ===========================================================================
C interface for tests (ctypes)
=========================================================================== */
extern "C" {

struct OrcLens
{
	double kappa, gamma, theta, kappastar;
	int rectangular;
	double corner_re, corner_im;
	int approx;
	int taylor_smooth;
};

struct OrcGrid
{
	double center_x[2];
	double half_length_x[2];
	double ray_half_sep[2];
	int num_ray_threads[2];   // cells along x1 and x2 (already doubled, ipm.cuh:468)
	double center_y[2];
	double half_length_y[2];
	int num_pixels[2];
};

struct OrcInput
{
	double kappa_tot, shear, kappa_star, theta_star;
	int mass_function;
	double m_solar, m_lower, m_upper, light_loss;
	int rectangular, approx;
	double safety_scale;
	double center_y[2];
	double half_length_y[2];
	int num_pixels_y[2];
	int num_rays_y;
};

struct OrcDerived
{
	double m_lower, m_upper;   // after m_solar scaling (ipm.cuh:381-390)
	double mean_mass, mean_mass2, mean_mass2_ln_mass;
	double mu_ave, num_rays_x;
	double ray_half_sep[2];
	double half_length_x[2];
	int num_ray_threads[2];
	double center_x[2];
	int num_stars;
	double corner[2];
	double alpha_error;
	int taylor_smooth;
	// the lens the geometry was derived for: the inputs, or the star file's
	// values (ipm.cuh:405-432) when the stars come from a file
	double kappa_star, theta_star;
	int rectangular;
	int covered;   // 0 when a star file's field is too small for the shooting region (ipm.cuh:521-529)
};

int orc_version() { return 1; }

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

void orc_star_uniforms(uint32_t seed, int64_t i, double* u4) { star_uniforms(seed, i, u4); }

void orc_mass_moments(int mf, double ml, double mu, double ms, double* out3) { mass_moments<double>(mf, ml, mu, ms, out3); }

double orc_mass_sample(int mf, double p, double ml, double mu, double ms) { return mass_sample(mf, p, ml, mu, ms); }

static void star_params(int n, int rectangular, double c_re, double c_im, double theta, const double* stars, double* out);

/* ipm.cuh:370-565. With file == nullptr the masses come from the mass
   function and the star count and field from the sizing formulas; otherwise
   from the star file (ipm.cuh:405-432, the moments of star.cuh:124-161 over
   the file's stars and corner), and a circular field is turned towards the
   shooting region's corner with the file's radius (ipm.cuh:507-529). */
struct OrcStarFile
{
	int n, rectangular;
	double corner[2];
	double theta;
	const double* stars;
};

/* The reference derives in its value type T (IPM<T>, ipm.cuh:370-565):
   lib_ipm.so in float, lib_ipm_double.so in double. derive<T> evaluates each
   formula in T with the reference's operand types (its inputs and members are
   T, the literals 1.0 and the pixel counts as written there), so a float
   evaluation lands on the same side of each ceil and floor as the float
   build. The mass moments come from mass_moments<T>; the moments of a star
   file accumulate in double over T masses as in star.cuh:124-161. Results are stored widened to double. */
}  // extern "C"

template <typename T>
static void star_params_t(int n, int rectangular, T c_re, T c_im, T theta, const double* stars, T* out)
{
	T mlo = std::numeric_limits<T>::max();
	T mhi = std::numeric_limits<T>::min();
	double mt = 0, m2 = 0, m2l = 0;
	for (int i = 0; i < n; i++)
	{
		T m = T(stars[3 * i + 2]);
		mt += m;
		m2 += m * m;
		m2l += m * m * std::log(m);
		mlo = std::min(mlo, m);
		mhi = std::max(mhi, m);
	}
	const T pi = T(PI);
	if (rectangular)
		out[0] = T(mt * pi * theta * theta / (4 * c_re * c_im));
	else
		out[0] = T(mt * theta * theta / (std::hypot(c_re, c_im) * std::hypot(c_re, c_im)));
	out[1] = mlo;
	out[2] = mhi;
	out[3] = T(mt / n);
	out[4] = T(m2 / n);
	out[5] = T(m2l / n);
}

template <typename T>
static void derive(const OrcInput* in, const OrcStarFile* file, OrcDerived* d)
{
	const T kt = T(in->kappa_tot), g = T(in->shear);
	T ks = T(in->kappa_star), th = T(in->theta_star);
	const T ll = T(in->light_loss), ss = T(in->safety_scale);
	const T pi = T(PI), inv_pi = T(1 / PI);
	int rect = in->rectangular;
	T mean_mass, mean_mass2;
	if (!file)
	{
		T ml = 1, mu = 1;
		if (in->mass_function != MF_EQUAL)
		{
			ml = T(in->m_lower) * T(in->m_solar);
			mu = T(in->m_upper) * T(in->m_solar);
		}
		d->m_lower = ml;
		d->m_upper = mu;
		T mm[3];
		mass_moments<T>(in->mass_function, ml, mu, T(in->m_solar), mm);
		mean_mass = mm[0];
		mean_mass2 = mm[1];
		d->mean_mass2_ln_mass = mm[2];
	}
	else
	{
		T sp[6];
		star_params_t<T>(file->n, file->rectangular, T(file->corner[0]), T(file->corner[1]), T(file->theta), file->stars,
			sp);
		rect = file->rectangular;
		th = T(file->theta);
		ks = sp[0];
		d->m_lower = sp[1];
		d->m_upper = sp[2];
		mean_mass = sp[3];
		mean_mass2 = sp[4];
		d->mean_mass2_ln_mass = sp[5];
	}
	d->mean_mass = mean_mass;
	d->mean_mass2 = mean_mass2;
	d->kappa_star = ks;
	d->theta_star = th;
	d->rectangular = rect;
	d->covered = 1;
	d->mu_ave = T(1 / ((1 - kt) * (1 - kt) - g * g));
	const T h1 = T(in->half_length_y[0]), h2 = T(in->half_length_y[1]);
	const int n1 = in->num_pixels_y[0], n2 = in->num_pixels_y[1];
	const T nrx = T(1.0 * in->num_rays_y * n1 * n2 / (2 * h1 * 2 * h2));
	d->num_rays_x = nrx;
	T rhs1 = std::sqrt(h1 / h2 * n2 / n1);
	T rhs2 = std::sqrt(h2 / h1 * n1 / n2);
	const T s = 2 * std::sqrt(nrx);
	rhs1 /= s;
	rhs2 /= s;
	d->ray_half_sep[0] = rhs1;
	d->ray_half_sep[1] = rhs2;
	const T ext = th * std::sqrt(ks * mean_mass2 / (mean_mass * ll));
	const T hx1 = (h1 + ext) / std::abs(1 - kt + g);
	const T hx2 = (h2 + ext) / std::abs(1 - kt - g);
	const int nrt1 = int(hx1 / (2 * rhs1)) + 1;
	const int nrt2 = int(hx2 / (2 * rhs2)) + 1;
	const T hl1 = 2 * rhs1 * nrt1, hl2 = 2 * rhs2 * nrt2;
	d->half_length_x[0] = hl1;
	d->half_length_x[1] = hl2;
	d->num_ray_threads[0] = 2 * nrt1;
	d->num_ray_threads[1] = 2 * nrt2;
	const T cx1 = T(in->center_y[0]) / (1 - kt + g);
	const T cx2 = T(in->center_y[1]) / (1 - kt - g);
	d->center_x[0] = cx1;
	d->center_x[1] = cx2;
	const T c1 = std::abs(cx1) + hl1;
	const T c2 = std::abs(cx2) + hl2;
	T k1, k2;   // the field's corner
	if (!file)
	{
		if (rect)
		{
			d->num_stars = int(std::ceil(ss * 2 * c1 * ss * 2 * c2 * ks / (pi * th * th * mean_mass)));
			const T a = std::sqrt(c1 / c2), b = std::sqrt(c2 / c1);
			const T f = std::sqrt(pi * th * th * d->num_stars * mean_mass / (4 * ks));
			k1 = a * f;
			k2 = b * f;
		}
		else
		{
			const T ca = std::hypot(c1, c2);
			d->num_stars = int(std::ceil(ss * ca * ss * ca * ks / (th * th * mean_mass)));
			const T f = std::sqrt(th * th * d->num_stars * mean_mass / ks);
			k1 = c1 / ca * f;
			k2 = c2 / ca * f;
		}
	}
	else
	{
		d->num_stars = file->n;
		k1 = T(file->corner[0]);
		k2 = T(file->corner[1]);
		const T ca = std::hypot(c1, c2);
		if (!rect)
		{
			const T r = std::hypot(k1, k2);
			k1 = c1 / ca * r;
			k2 = c2 / ca * r;
		}
		if ((rect && (k1 < ss * c1 || k2 < ss * c2)) || (!rect && std::hypot(k1, k2) < ss * ca))
			d->covered = 0;
	}
	d->corner[0] = k1;
	d->corner[1] = k2;
	const T ae = std::min(h1 / (10 * n1), h2 / (10 * n2));
	d->alpha_error = ae;
	int ts = 1;
	const T cabs_ = std::hypot(k1, k2);
	const T carg_ = std::atan2(k2, k1);
	while ((ks * inv_pi * 4 / (ts + 1) * cabs_ * (ss + 1) / (ss - 1) * std::pow(1 / ss, ts + 1) > ae) && ts <= 101)
		ts += 2;
	while ((std::fmod(carg_ * (ts - 1), pi) < 0.1 * pi || std::fmod(carg_ * (ts - 1), pi) > 0.9 * pi) && ts <= 101)
		ts += 2;
	d->taylor_smooth = ts;
}

extern "C" {

/* single = 1: evaluated in float, as the float library derives */
int orc_derived_params(const OrcInput* in, int single, OrcDerived* d)
{
	if (single)
		derive<float>(in, nullptr, d);
	else
		derive<double>(in, nullptr, d);
	return 0;
}

/* the star-file branch: n stars (AoS x1, x2, m) with the file's field shape,
   corner and theta; returns 0, or -1 when the field does not cover the
   shooting region (the reference's run then fails, ipm.cuh:521-529) */
int orc_derived_params_file(const OrcInput* in, int single, int n, int rectangular, double c_re, double c_im,
	double theta, const double* stars, OrcDerived* d)
{
	OrcStarFile f{n, rectangular, {c_re, c_im}, theta, stars};
	if (single)
		derive<float>(in, &f, d);
	else
		derive<double>(in, &f, d);
	return d->covered ? 0 : -1;
}

/* the corner after the field is realised (ipm.cuh:717-731): the derived
   corner's shape with the size set by the realised mean mass; the stars do
   not move. out = [corner_re, corner_im, root_half_length (ipm.cuh:746-754)] */
}  // extern "C"

template <typename T>
static void rescaled_corner(int rectangular, T c_re, T c_im, T theta, int n, T mean_mass_actual, T kappa_star, double* out)
{
	T r, i;
	if (rectangular)
	{
		const T f = std::sqrt(T(PI) * theta * theta * n * mean_mass_actual / (4 * kappa_star));
		r = std::sqrt(c_re / c_im) * f;
		i = std::sqrt(c_im / c_re) * f;
		out[2] = T((r > i ? r : i) * 1.1);
	}
	else
	{
		const T a = std::hypot(c_re, c_im);
		const T f = std::sqrt(theta * theta * n * mean_mass_actual / kappa_star);
		r = c_re / a * f;
		i = c_im / a * f;
		out[2] = T(std::hypot(r, i) * 1.1);
	}
	out[0] = r;
	out[1] = i;
}

extern "C" {

/* the tree's root half-length for a field corner (ipm.cuh:746-754): the larger
   side of a rectangle or the radius of a disk, times 1.1, rounded to T */
double orc_root_half_length(int rectangular, int single, double c_re, double c_im)
{
	if (single)
	{
		const float r = float(c_re), i = float(c_im);
		return float((rectangular ? (r > i ? r : i) : std::hypot(r, i)) * 1.1);
	}
	return (rectangular ? (c_re > c_im ? c_re : c_im) : std::hypot(c_re, c_im)) * 1.1;
}

void orc_rescaled_corner(int rectangular, int single, double c_re, double c_im, double theta, int n,
	double mean_mass_actual, double kappa_star, double* out)
{
	if (single)
		rescaled_corner<float>(rectangular, float(c_re), float(c_im), float(theta), n, float(mean_mass_actual),
			float(kappa_star), out);
	else
		rescaled_corner<double>(rectangular, c_re, c_im, theta, n, mean_mass_actual, kappa_star, out);
}

// star.cuh:64-107 with the Philox stream above; out = AoS [x1, x2, m] * n
void orc_generate_stars(int n, int rectangular, double c_re, double c_im, int mf, double ml, double mu, double ms,
	uint32_t seed, double* out)
{
#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++)
	{
		double u[4];
		star_uniforms(seed, i, u);
		double x1, x2;
		if (rectangular)
		{
			x1 = u[0] * 2 * c_re - c_re;
			x2 = u[1] * 2 * c_im - c_im;
		}
		else
		{
			double a = u[0] * 2 * PI;
			double r = std::sqrt(u[1]) * std::hypot(c_re, c_im);
			x1 = r * std::cos(a);
			x2 = r * std::sin(a);
		}
		out[3 * i + 0] = x1;
		out[3 * i + 1] = x2;
		out[3 * i + 2] = mass_sample(mf, u[2], ml, mu, ms);
	}
}

// star.cuh:124-161; out = [kappastar, m_lower, m_upper, mean_mass, mean_mass2, mean_mass2_ln_mass]
void orc_star_params(int n, int rectangular, double c_re, double c_im, double theta, const double* stars, double* out)
{
	star_params(n, rectangular, c_re, c_im, theta, stars, out);
}

static void star_params(int n, int rectangular, double c_re, double c_im, double theta, const double* stars, double* out)
{
	double mlo = std::numeric_limits<double>::max();
	double mhi = std::numeric_limits<double>::min();
	double mt = 0, m2 = 0, m2l = 0;
	for (int i = 0; i < n; i++)
	{
		double m = stars[3 * i + 2];
		mt += m;
		m2 += m * m;
		m2l += m * m * std::log(m);
		mlo = std::min(mlo, m);
		mhi = std::max(mhi, m);
	}
	if (rectangular)
		out[0] = mt * PI * theta * theta / (4 * c_re * c_im);
	else
		out[0] = mt * theta * theta / (std::hypot(c_re, c_im) * std::hypot(c_re, c_im));
	out[1] = mlo;
	out[2] = mhi;
	out[3] = mt / n;
	out[4] = m2 / n;
	out[5] = m2l / n;
}

void* orc_tree_build(const double* stars, int n, double root_half_length)
{
	Tree* t = new Tree();
	t->sx.resize(n);
	t->sy.resize(n);
	t->sm.resize(n);
	for (int i = 0; i < n; i++)
	{
		t->sx[i] = stars[3 * i];
		t->sy[i] = stars[3 * i + 1];
		t->sm[i] = stars[3 * i + 2];
	}
	build_tree(*t, root_half_length);
	return t;
}

void orc_tree_free(void* h) { delete static_cast<Tree*>(h); }
int orc_tree_levels(void* h) { return static_cast<Tree*>(h)->levels; }
int orc_tree_num_nodes(void* h) { return int(static_cast<Tree*>(h)->nodes.size()); }

void orc_tree_stars(void* h, double* out)
{
	Tree* t = static_cast<Tree*>(h);
	for (size_t i = 0; i < t->sx.size(); i++)
	{
		out[3 * i] = t->sx[i];
		out[3 * i + 1] = t->sy[i];
		out[3 * i + 2] = t->sm[i];
	}
}

/* node table: geom = [cx, cy, hl] per node; info = [level, parent, child0, nchild, begin, n, nnbr, nm2l, np2l];
   lists = [nbr x8, m2l x27, p2l x5] per node (-1 padded) */
void orc_tree_export(void* h, double* geom, int* info, int* lists)
{
	Tree* t = static_cast<Tree*>(h);
	for (size_t i = 0; i < t->nodes.size(); i++)
	{
		const Node& N = t->nodes[i];
		geom[3 * i] = N.cx;
		geom[3 * i + 1] = N.cy;
		geom[3 * i + 2] = N.hl;
		int* f = info + 9 * i;
		f[0] = N.level; f[1] = N.parent; f[2] = N.child0; f[3] = N.nchild; f[4] = N.begin; f[5] = N.n;
		f[6] = N.nnbr; f[7] = N.nm2l; f[8] = N.np2l;
		int* L = lists + 40 * i;
		for (int j = 0; j < 40; j++) L[j] = -1;
		for (int j = 0; j < N.nnbr; j++) L[j] = N.nbr[j];
		for (int j = 0; j < N.nm2l; j++) L[8 + j] = N.m2l[j];
		for (int j = 0; j < N.np2l; j++) L[35 + j] = N.p2l[j];
	}
}

// ipm.cuh:848-860; returns the automatic order before clamping
int orc_expansion_order(double theta, int levels, double root_hl, double alpha_error)
{
	return int(std::ceil(2 * std::log2(theta) + levels - std::log2(root_hl) - std::log2(alpha_error)));
}

int orc_tree_fmm(void* h, int p)
{
	if (p < 1 || p > MAX_EXPANSION_ORDER) return -1;
	fmm(*static_cast<Tree*>(h), p);
	return 0;
}

/* coefficients: out = 2*(p+1) doubles per node (re, im) */
void orc_tree_coeffs(void* h, int which, double* out)
{
	Tree* t = static_cast<Tree*>(h);
	const std::vector<Cx>& v = which ? t->local : t->mult;
	for (size_t i = 0; i < v.size(); i++)
	{
		out[2 * i] = v[i].re;
		out[2 * i + 1] = v[i].im;
	}
}

/* w = lens equation at z via the tree (ipm_functions.cuh:34-46); mu optional */
void orc_deflect_tree(void* h, const OrcLens* Lp, const double* z, int n, double* w, double* mu, int* leaf)
{
	Tree* t = static_cast<Tree*>(h);
	Lens L{Lp->kappa, Lp->gamma, Lp->theta, Lp->kappastar, Lp->rectangular, Lp->corner_re, Lp->corner_im, Lp->approx, Lp->taylor_smooth};
#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++)
	{
		int lf;
		DeflectOut o = deflect_tree(*t, L, cx(z[2 * i], z[2 * i + 1]), mu != nullptr, &lf);
		w[2 * i] = o.w.re;
		w[2 * i + 1] = o.w.im;
		if (mu) mu[i] = o.mu;
		if (leaf) leaf[i] = lf;
	}
}

/* w = lens equation at z by direct summation over every star (no tree) */
void orc_deflect_direct(const double* stars, int nstars, const OrcLens* Lp, const double* z, int n, double* w)
{
	Lens L{Lp->kappa, Lp->gamma, Lp->theta, Lp->kappastar, Lp->rectangular, Lp->corner_re, Lp->corner_im, Lp->approx, Lp->taylor_smooth};
#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++)
	{
		Cx zz = cx(z[2 * i], z[2 * i + 1]);
		Cx as = cx(0);
		for (int s = 0; s < nstars; s++) as = as + cx(stars[3 * s + 2]) / (zz - cx(stars[3 * s], stars[3 * s + 1]));
		as = conj(as * (L.theta * L.theta));
		Cx o = (1 - L.kappa) * zz + L.gamma * conj(zz) - as - alpha_smooth(zz, L);
		w[2 * i] = o.re;
		w[2 * i + 1] = o.im;
	}
}

/* one triangle (pixel coordinates, already flipped) through polygon.cuh:349-467.
   Returns the number of contributions written (capped at cap). */
int orc_triangle_alloc(const double* tri, double factor, int n1, int n2, long long* idx, double* val, int cap)
{
	Poly P;
	for (int k = 0; k < 3; k++) P.add(cx(tri[2 * k], tri[2 * k + 1]));
	int cnt = 0;
	allocate_area_among_pixels(P, factor, n1, n2, [&](size_t i, double v) {
		if (cnt < cap)
		{
			idx[cnt] = (long long)i;
			val[cnt] = v;
		}
		cnt++;
	});
	return cnt;
}

/* Corner lattice point (I, J), I in [0, N1], J in [0, N2]: the corners of the
   reference's cell (i, j) (ipm_functions.cuh:125-130) are (i+1,j+1), (i,j+1),
   (i,j), (i+1,j). */
static inline Cx lattice_corner(const OrcGrid& G, long long I, long long J)
{
	return cx(G.center_x[0] - G.half_length_x[0] + 2 * G.ray_half_sep[0] * double(I),
		G.center_x[1] - G.half_length_x[1] + 2 * G.ray_half_sep[1] * double(J));
}

/* source pixel coordinates of a corner: ipm_functions.cuh:159-165 */
static inline Cx to_pixel(const OrcGrid& G, Cx w)
{
	w = w - cx(G.center_y[0], G.center_y[1]);
	Cx y = cx((w.re + G.half_length_y[0]) * G.num_pixels[0] / (2 * G.half_length_y[0]),
		(w.im + G.half_length_y[1]) * G.num_pixels[1] / (2 * G.half_length_y[1]));
	y.im = G.num_pixels[1] - y.im;
	return y;
}

/* IPM: shoot cell rows rows[0..nrows) of the lattice (ipm_functions.cuh:112-230)
   and add their areas into map (n2 x n1, row 0 at the top). If minima/saddles
   are given, triangles go to them by parity instead (ipm_functions.cuh:178-191).
   use_direct != 0 deflects by direct summation over all stars instead of the tree. */
void orc_ipm_rows(void* h, const OrcLens* Lp, const OrcGrid* Gp, const long long* rows, int nrows, int use_direct,
	double* map, double* minima, double* saddles)
{
	Tree* t = static_cast<Tree*>(h);
	const OrcGrid G = *Gp;
	Lens L{Lp->kappa, Lp->gamma, Lp->theta, Lp->kappastar, Lp->rectangular, Lp->corner_re, Lp->corner_im, Lp->approx, Lp->taylor_smooth};
	const long long N1 = G.num_ray_threads[0];
	const int n1 = G.num_pixels[0], n2 = G.num_pixels[1];
	const double factor = 2 * G.ray_half_sep[0] * G.ray_half_sep[1] * n1 * n2 / (2 * G.half_length_y[0] * 2 * G.half_length_y[1]);
	const double th2 = L.theta * L.theta;
	auto deflect = [&](Cx z) -> Cx {
		if (!use_direct) return deflect_tree(*t, L, z, false).w;
		Cx as = cx(0);
		for (size_t s = 0; s < t->sx.size(); s++) as = as + cx(t->sm[s]) / (z - cx(t->sx[s], t->sy[s]));
		as = conj(as * th2);
		return (1 - L.kappa) * z + L.gamma * conj(z) - as - alpha_smooth(z, L);
	};
#pragma omp parallel
	{
		std::vector<Cx> lo(N1 + 1), hi(N1 + 1);
#pragma omp for schedule(dynamic, 1)
		for (int r = 0; r < nrows; r++)
		{
			long long j = rows[r];
			for (long long I = 0; I <= N1; I++)
			{
				lo[I] = to_pixel(G, deflect(lattice_corner(G, I, j)));
				hi[I] = to_pixel(G, deflect(lattice_corner(G, I, j + 1)));
			}
			for (long long i = 0; i < N1; i++)
			{
				Cx y0 = hi[i + 1], y1 = hi[i], y2 = lo[i], y3 = lo[i + 1];
				bool bad[4];
				Cx yy[4] = {y0, y1, y2, y3};
				for (int a = 0; a < 4; a++) bad[a] = std::isnan(yy[a].re) || std::isnan(yy[a].im);
				const int tris[2][3] = {{0, 1, 2}, {2, 3, 0}};
				for (int k = 0; k < 2; k++)
				{
					if (bad[tris[k][0]] || bad[tris[k][1]] || bad[tris[k][2]]) continue;
					Poly P;
					P.add(yy[tris[k][0]]);
					P.add(yy[tris[k][1]]);
					P.add(yy[tris[k][2]]);
					double A = P.area();
					if (!(std::fabs(A) < 10000 * factor)) continue;
					double* dst = map;
					if (minima && saddles) dst = (A < 0) ? minima : saddles;
					allocate_area_among_pixels(P, factor, n1, n2, [&](size_t idx, double v) { atomic_add(&dst[idx], v); });
				}
			}
		}
	}
}

/* IPM over the cells named by window records (lattice row, first column of a
   63-cell tile, bit mask of its cells), keeping only the area that lands in map
   columns [wx0, wx0 + ww) and rows [wy0, wy0 + wh); window is wh x ww, its row 0
   the map's row wy0. The triangles and their allocation are those of
   orc_ipm_rows; the corners come from the tree (FP64). */
void orc_ipm_window(void* h, const OrcLens* Lp, const OrcGrid* Gp, const long long* rec, long long nrec, int wx0, int wy0,
	int ww, int wh, double* window)
{
	Tree* t = static_cast<Tree*>(h);
	const OrcGrid G = *Gp;
	Lens L{Lp->kappa, Lp->gamma, Lp->theta, Lp->kappastar, Lp->rectangular, Lp->corner_re, Lp->corner_im, Lp->approx, Lp->taylor_smooth};
	const long long N1 = G.num_ray_threads[0];
	const int n1 = G.num_pixels[0], n2 = G.num_pixels[1];
	const double factor = 2 * G.ray_half_sep[0] * G.ray_half_sep[1] * n1 * n2 / (2 * G.half_length_y[0] * 2 * G.half_length_y[1]);
#pragma omp parallel
	{
		Cx lo[64], hi[64];
#pragma omp for schedule(dynamic, 64)
		for (long long r = 0; r < nrec; r++)
		{
			const long long j = rec[3 * r], i0 = rec[3 * r + 1];
			const unsigned long long mask = (unsigned long long)rec[3 * r + 2];
			for (int b = 0; b < 64; b++)
			{
				const bool used = ((mask >> b) & 1ull) || (b > 0 && ((mask >> (b - 1)) & 1ull));
				if (!used || i0 + b > N1) continue;
				lo[b] = to_pixel(G, deflect_tree(*t, L, lattice_corner(G, i0 + b, j), false).w);
				hi[b] = to_pixel(G, deflect_tree(*t, L, lattice_corner(G, i0 + b, j + 1), false).w);
			}
			for (int b = 0; b < 63; b++)
			{
				if (!((mask >> b) & 1ull) || i0 + b >= N1) continue;
				Cx yy[4] = {hi[b + 1], hi[b], lo[b], lo[b + 1]};
				bool bad[4];
				for (int a = 0; a < 4; a++) bad[a] = std::isnan(yy[a].re) || std::isnan(yy[a].im);
				const int tris[2][3] = {{0, 1, 2}, {2, 3, 0}};
				for (int k = 0; k < 2; k++)
				{
					if (bad[tris[k][0]] || bad[tris[k][1]] || bad[tris[k][2]]) continue;
					Poly P;
					P.add(yy[tris[k][0]]);
					P.add(yy[tris[k][1]]);
					P.add(yy[tris[k][2]]);
					double A = P.area();
					if (!(std::fabs(A) < 10000 * factor)) continue;
					allocate_area_among_pixels(P, factor, n1, n2, [&](size_t idx, double v) {
						const long long y = (long long)(idx / size_t(n1)), x = (long long)(idx % size_t(n1));
						if (x < wx0 || x >= wx0 + ww || y < wy0 || y >= wy0 + wh) return;
						atomic_add(&window[(y - wy0) * ww + (x - wx0)], v);
					});
				}
			}
		}
	}
}

/* IRS: shoot ray rows (irs_functions.cuh:111-212) into int counts; returns
   the number of rays received. Parity maps use the sign of mu. */
long long orc_irs_rows(void* h, const OrcLens* Lp, const OrcGrid* Gp, const long long* rows, int nrows,
	int* map, int* minima, int* saddles)
{
	Tree* t = static_cast<Tree*>(h);
	const OrcGrid G = *Gp;
	Lens L{Lp->kappa, Lp->gamma, Lp->theta, Lp->kappastar, Lp->rectangular, Lp->corner_re, Lp->corner_im, Lp->approx, Lp->taylor_smooth};
	const long long N1 = G.num_ray_threads[0];
	const int n1 = G.num_pixels[0], n2 = G.num_pixels[1];
	long long received = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : received)
	for (int r = 0; r < nrows; r++)
	{
		long long j = rows[r];
		for (long long i = 0; i < N1; i++)
		{
			Cx z = cx(G.center_x[0] - G.half_length_x[0] + G.ray_half_sep[0] + 2 * G.ray_half_sep[0] * double(i),
				G.center_x[1] - G.half_length_x[1] + G.ray_half_sep[1] + 2 * G.ray_half_sep[1] * double(j));
			bool want_mu = minima && saddles;
			DeflectOut o = deflect_tree(*t, L, z, want_mu);
			if (std::isnan(o.w.re) || std::isnan(o.w.im)) continue;
			Cx w = o.w - cx(G.center_y[0], G.center_y[1]);
			if (w.re < -G.half_length_y[0] || w.re > G.half_length_y[0] || w.im < -G.half_length_y[1] || w.im > G.half_length_y[1]) continue;
			int px = int((w.re + G.half_length_y[0]) * n1 / (2 * G.half_length_y[0]));
			int py = int((w.im + G.half_length_y[1]) * n2 / (2 * G.half_length_y[1]));
			if (px == n1) px--;
			if (py == n2) py--;
			py = n2 - 1 - py;
			size_t idx = size_t(py) * n1 + px;
			if (want_mu)
			{
				int* dst = (o.mu >= 0) ? minima : saddles;
#pragma omp atomic
				dst[idx] += 1;
			}
#pragma omp atomic
			map[idx] += 1;
			received++;
		}
	}
	return received;
}

/* IRS pixel of a source-plane position w (already minus center_y), exactly as
   irs_functions.cuh:152-184: -1 when outside the map */
long long orc_irs_bin(double w1, double w2, double h1, double h2, int n1, int n2)
{
	if (w1 < -h1 || w1 > h1 || w2 < -h2 || w2 > h2) return -1;
	int px = int((w1 + h1) * n1 / (2 * h1));
	int py = int((w2 + h2) * n2 / (2 * h2));
	if (px == n1) px--;
	if (py == n2) py--;
	py = n2 - 1 - py;
	return (long long)py * n1 + px;
}

int orc_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
