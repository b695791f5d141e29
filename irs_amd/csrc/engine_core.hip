// Engine: device buffers, star field, quadtree build, FMM and leaf tables.
#include "engine.hpp"
#include "device_math.hpp"
#include "mass_functions.hpp"
#include "philox.hpp"

#include <algorithm>
#include <cstring>

namespace mlens
{

using dev::dcx;

template <typename T>
void DeviceBuffer<T>::alloc(size_t n)
{
	if (n <= count && ptr) return;
	free();
	if (n == 0) return;
	MLENS_HIP_CHECK(hipMalloc(&ptr, n * sizeof(T)));
	count = n;
}

template <typename T>
void DeviceBuffer<T>::free()
{
	if (ptr) (void)hipFree(ptr);
	ptr = nullptr;
	count = 0;
}

template struct DeviceBuffer<double>;
template struct DeviceBuffer<float>;
template struct DeviceBuffer<int>;
template struct DeviceBuffer<char>;
template struct DeviceBuffer<long long>;
template struct DeviceBuffer<unsigned long long>;
template struct DeviceBuffer<unsigned short>;

static inline unsigned blocks_for(long long n, int threads) { return unsigned((n + threads - 1) / threads); }

Engine::Engine()
{
	MLENS_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
	MLENS_HIP_CHECK(hipEventCreate(&ev0_));
	MLENS_HIP_CHECK(hipEventCreate(&ev1_));
	int dev = 0;
	MLENS_HIP_CHECK(hipGetDevice(&dev));
	dev_ = dev;
	MLENS_HIP_CHECK(hipDeviceGetAttribute(&num_cus_, hipDeviceAttributeMultiprocessorCount, dev));
}

Engine::~Engine()
{
	if (stream_) (void)hipStreamSynchronize(stream_);
	if (ev0_) (void)hipEventDestroy(ev0_);
	if (ev1_) (void)hipEventDestroy(ev1_);
	for (hipEvent_t e : m2l_ev_) (void)hipEventDestroy(e);
	if (stream_) (void)hipStreamDestroy(stream_);
}

/* ---------------------------------------------------------------------------
Stars
--------------------------------------------------------------------------- */
void Engine::upload_stars(const double* aos, int n)
{
	nstars_ = n;
	sx_.alloc(n);
	sy_.alloc(n);
	sm_.alloc(n);
	std::vector<double> a(n), b(n), c(n);
	for (int i = 0; i < n; i++)
	{
		a[i] = aos[3 * i];
		b[i] = aos[3 * i + 1];
		c[i] = aos[3 * i + 2];
	}
	MLENS_HIP_CHECK(hipMemcpy(sx_.ptr, a.data(), n * sizeof(double), hipMemcpyHostToDevice));
	MLENS_HIP_CHECK(hipMemcpy(sy_.ptr, b.data(), n * sizeof(double), hipMemcpyHostToDevice));
	MLENS_HIP_CHECK(hipMemcpy(sm_.ptr, c.data(), n * sizeof(double), hipMemcpyHostToDevice));
}

// star.cuh:64-107 with the Philox stream of philox.hpp
__global__ void k_generate_stars(double* sx, double* sy, double* sm, int n, int rectangular, double c_re, double c_im,
	int mf, double ml, double mu, double ms, uint32_t seed)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
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
		double r = sqrt(u[1]) * hypot(c_re, c_im);
		x1 = r * cos(a);
		x2 = r * sin(a);
	}
	sx[i] = x1;
	sy[i] = x2;
	sm[i] = sample_mass<double>(mf, u[2], ml, mu, ms);
}

void Engine::generate_stars(int n, int rectangular, double c_re, double c_im, int mf, double ml, double mu, double ms,
	uint32_t seed)
{
	nstars_ = n;
	sx_.alloc(n);
	sy_.alloc(n);
	sm_.alloc(n);
	k_generate_stars<<<blocks_for(n, 256), 256, 0, stream_>>>(sx_.ptr, sy_.ptr, sm_.ptr, n, rectangular, c_re, c_im, mf,
		ml, mu, ms, seed);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::download_stars(double* aos) const
{
	int n = nstars_;
	std::vector<double> a(n), b(n), c(n);
	MLENS_HIP_CHECK(hipMemcpy(a.data(), sx_.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
	MLENS_HIP_CHECK(hipMemcpy(b.data(), sy_.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
	MLENS_HIP_CHECK(hipMemcpy(c.data(), sm_.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
	for (int i = 0; i < n; i++)
	{
		aos[3 * i] = a[i];
		aos[3 * i + 1] = b[i];
		aos[3 * i + 2] = c[i];
	}
}

// star.cuh:124-161: totals in FP64; one block reduces everything (N fits easily)
__global__ void k_star_stats(const double* sm, int n, double* out)
{
	__shared__ double s_t[256], s_2[256], s_l[256], s_lo[256], s_hi[256];
	double t = 0, t2 = 0, tl = 0, lo = 1e300, hi = -1e300;
	for (int i = threadIdx.x; i < n; i += blockDim.x)
	{
		double m = sm[i];
		t += m;
		t2 += m * m;
		tl += m * m * log(m);
		lo = fmin(lo, m);
		hi = fmax(hi, m);
	}
	int k = threadIdx.x;
	s_t[k] = t;
	s_2[k] = t2;
	s_l[k] = tl;
	s_lo[k] = lo;
	s_hi[k] = hi;
	__syncthreads();
	for (int off = blockDim.x / 2; off > 0; off >>= 1)
	{
		if (k < off)
		{
			s_t[k] += s_t[k + off];
			s_2[k] += s_2[k + off];
			s_l[k] += s_l[k + off];
			s_lo[k] = fmin(s_lo[k], s_lo[k + off]);
			s_hi[k] = fmax(s_hi[k], s_hi[k + off]);
		}
		__syncthreads();
	}
	if (k == 0)
	{
		out[0] = s_t[0];
		out[1] = s_2[0];
		out[2] = s_l[0];
		out[3] = s_lo[0];
		out[4] = s_hi[0];
	}
}

StarStats Engine::star_stats()
{
	scratch_d_.alloc(8);
	k_star_stats<<<1, 256, 0, stream_>>>(sm_.ptr, nstars_, scratch_d_.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	double h[5];
	MLENS_HIP_CHECK(hipMemcpyAsync(h, scratch_d_.ptr, sizeof(h), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	return StarStats{h[0], h[1], h[2], h[3], h[4]};
}

/* ---------------------------------------------------------------------------
Exclusive scan of int4 records (three passes: tiles, tile totals, add-back).
out has n + 1 entries; out[n] is the grand total.
--------------------------------------------------------------------------- */
struct I4
{
	int x, y, z, w;
};
__device__ __forceinline__ I4 operator+(I4 a, I4 b) { return I4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ I4 operator-(I4 a, I4 b) { return I4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
__device__ __forceinline__ int comp(I4 a, int q) { return q == 0 ? a.x : (q == 1 ? a.y : (q == 2 ? a.z : a.w)); }

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ inline I4 block_exclusive_scan(I4 v, I4* s, I4& total)
{
	int t = threadIdx.x;
	s[t] = v;
	__syncthreads();
	for (int off = 1; off < SCAN_THREADS; off <<= 1)
	{
		I4 add = t >= off ? s[t - off] : I4{0, 0, 0, 0};
		__syncthreads();
		s[t] = s[t] + add;
		__syncthreads();
	}
	total = s[SCAN_THREADS - 1];
	I4 ex = t > 0 ? s[t - 1] : I4{0, 0, 0, 0};
	__syncthreads();
	return ex;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const I4* in, I4* out, I4* tile_sums, int n)
{
	__shared__ I4 s[SCAN_THREADS];
	long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
	I4 v[SCAN_ITEMS];
	I4 sum{0, 0, 0, 0};
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++)
	{
		v[k] = (base + k < n) ? in[base + k] : I4{0, 0, 0, 0};
		sum = sum + v[k];
	}
	I4 total;
	I4 run = block_exclusive_scan(sum, s, total);
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++)
	{
		if (base + k < n) out[base + k] = run;
		run = run + v[k];
	}
	if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums(I4* tile_sums, int ntiles, I4* grand)
{
	__shared__ I4 s[SCAN_THREADS];
	I4 carry{0, 0, 0, 0};
	for (int base = 0; base < ntiles; base += SCAN_THREADS)
	{
		int i = base + threadIdx.x;
		I4 v = i < ntiles ? tile_sums[i] : I4{0, 0, 0, 0};
		I4 total;
		I4 ex = block_exclusive_scan(v, s, total);
		if (i < ntiles) tile_sums[i] = ex + carry;
		carry = carry + total;
	}
	if (threadIdx.x == 0) *grand = carry;
}

__global__ void k_scan_add(I4* out, const I4* tile_sums, int n, const I4* grand)
{
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = out[i] + tile_sums[i / SCAN_TILE];
	if (i == n) out[n] = *grand;
}

static void scan_i4(const I4* in, I4* out, I4* tiles, int n, hipStream_t st)
{
	int ntiles = std::max(1, (n + SCAN_TILE - 1) / SCAN_TILE);
	k_scan_tiles<<<ntiles, SCAN_THREADS, 0, st>>>(in, out, tiles, n);
	k_scan_tile_sums<<<1, SCAN_THREADS, 0, st>>>(tiles, ntiles, tiles + ntiles);
	k_scan_add<<<blocks_for(n + 1, 256), 256, 0, st>>>(out, tiles, n, tiles + ntiles);
	MLENS_HIP_CHECK(hipGetLastError());
}

/* ---------------------------------------------------------------------------
Quadtree (tree_node.cuh:32-561, level loop of ipm.cuh:740-839)
--------------------------------------------------------------------------- */
struct TreeView
{
	double *cx, *cy, *hl;
	int *level, *parent, *child0, *begin, *count, *nnbr, *nm2l, *np2l, *nbr, *m2l, *p2l;
};

__device__ __forceinline__ int quadrant(double x, double y, double cx, double cy)   // tree_node.cuh:432-477
{
	if (y >= cy) return (x >= cx) ? 0 : 1;
	return (x < cx) ? 2 : 3;
}

// tree_node.cuh:321-345: stars in a node plus its neighbours decide subdivision
__global__ void k_near_count(TreeView t, int l0, int l1, I4* flags, int* maxcnt)
{
	int i = l0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= l1) return;
	int c = t.count[i];
	for (int j = 0; j < t.nnbr[i]; j++) c += t.count[t.nbr[8 * i + j]];
	flags[i - l0] = I4{c > MAX_NUM_STARS_DIRECT ? 1 : 0, 0, 0, 0};
	atomicMax(maxcnt, c);
}

// tree_node.cuh:151-177: four children, offsets (+,+), (-,+), (-,-), (+,-)
__global__ void k_make_children(TreeView t, int l0, int l1, const I4* flags, const I4* rank, int base)
{
	int i = l0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= l1 || !flags[i - l0].x) return;
	int c0 = base + 4 * rank[i - l0].x;
	t.child0[i] = c0;
	double nh = t.hl[i] / 2;
	const double ox[4] = {1, -1, -1, 1};
	const double oy[4] = {1, 1, -1, -1};
	for (int q = 0; q < 4; q++)
	{
		int c = c0 + q;
		t.cx[c] = t.cx[i] + nh * ox[q];
		t.cy[c] = t.cy[i] + nh * oy[q];
		t.hl[c] = nh;
		t.level[c] = t.level[i] + 1;
		t.parent[c] = i;
		t.child0[c] = -1;
		t.nnbr[c] = 0;
		t.nm2l[c] = 0;
		t.np2l[c] = 0;
	}
}

__global__ void k_star_onehot(TreeView t, const double* sx, const double* sy, const int* star_node, int l0,
	const I4* flags, I4* onehot, int n)
{
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	int node = star_node[s];
	I4 v{0, 0, 0, 0};
	if (node >= 0 && flags[node - l0].x)
	{
		int q = quadrant(sx[s], sy[s], t.cx[node], t.cy[node]);
		v.x = q == 0;
		v.y = q == 1;
		v.z = q == 2;
		v.w = q == 3;
	}
	onehot[s] = v;
}

__global__ void k_child_ranges(TreeView t, int l0, int l1, const I4* flags, const I4* pre)
{
	int i = l0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= l1 || !flags[i - l0].x) return;
	int b = t.begin[i];
	I4 c = pre[b + t.count[i]] - pre[b];
	int c0 = t.child0[i];
	int start = b;
	for (int q = 0; q < 4; q++)
	{
		int cq = comp(c, q);
		t.begin[c0 + q] = start;
		t.count[c0 + q] = cq;
		start += cq;
	}
}

// stable partition of each split node's stars into its four children (sort_stars_kernel, tree_node.cuh:392-501)
__global__ void k_star_scatter(TreeView t, const double* sx, const double* sy, const double* sm, double* tx, double* ty,
	double* tm, const int* star_node, int* star_node_out, int l0, const I4* flags, const I4* pre, int n)
{
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	int node = star_node[s];
	if (node >= 0 && flags[node - l0].x)
	{
		int q = quadrant(sx[s], sy[s], t.cx[node], t.cy[node]);
		int b = t.begin[node];
		int rank = comp(pre[s], q) - comp(pre[b], q);
		int child = t.child0[node] + q;
		int dst = t.begin[child] + rank;
		tx[dst] = sx[s];
		ty[dst] = sy[s];
		tm[dst] = sm[s];
		star_node_out[dst] = child;
	}
	else
	{
		tx[s] = sx[s];
		ty[s] = sy[s];
		tm[s] = sm[s];
		star_node_out[s] = -1;   // star now sits in a leaf
	}
}

// tree_node.cuh:183-267
__global__ void k_set_neighbors(TreeView t, int n0, int n1)
{
	int i = n0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n1) return;
	int par = t.parent[i];
	double cx = t.cx[i], cy = t.cy[i], h = t.hl[i];
	int nn = 0, nm = 0, np = 0;
	int pc0 = t.child0[par];
	for (int q = 0; q < 4; q++)
	{
		int c = pc0 + q;
		if (t.count[c] == 0 || c == i) continue;
		t.nbr[8 * i + nn++] = c;
	}
	for (int j = 0; j < t.nnbr[par]; j++)
	{
		int nb = t.nbr[8 * par + j];
		if (t.count[nb] == 0) continue;
		if (t.child0[nb] < 0)
		{
			if (fabs(t.cx[nb] - cx) < 3.5 * h && fabs(t.cy[nb] - cy) < 3.5 * h)
				t.nbr[8 * i + nn++] = nb;
			else
				t.p2l[5 * i + np++] = nb;
		}
		else
		{
			int c0 = t.child0[nb];
			for (int q = 0; q < 4; q++)
			{
				int c = c0 + q;
				if (t.count[c] == 0) continue;
				if (fabs(t.cx[c] - cx) < 3 * h && fabs(t.cy[c] - cy) < 3 * h)
					t.nbr[8 * i + nn++] = c;
				else
					t.m2l[27 * i + nm++] = c;
			}
		}
	}
	t.nnbr[i] = nn;
	t.nm2l[i] = nm;
	t.np2l[i] = np;
}

template <typename T>
static void grow(DeviceBuffer<T>& b, size_t used, size_t n)
{
	if (n <= b.count) return;
	T* p = nullptr;
	MLENS_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
	if (used) MLENS_HIP_CHECK(hipMemcpy(p, b.ptr, used * sizeof(T), hipMemcpyDeviceToDevice));
	b.free();
	b.ptr = p;
	b.count = n;
}

void Engine::ensure_node_capacity(int n)
{
	if (n <= node_cap_) return;
	int cap = std::max(n, std::max(1024, 2 * node_cap_));
	size_t u = size_t(num_nodes_);
	grow(ncx_, u, cap);
	grow(ncy_, u, cap);
	grow(nhl_, u, cap);
	grow(nlevel_, u, cap);
	grow(nparent_, u, cap);
	grow(nchild0_, u, cap);
	grow(nbegin_, u, cap);
	grow(ncount_, u, cap);
	grow(nnnbr_, u, cap);
	grow(nnm2l_, u, cap);
	grow(nnp2l_, u, cap);
	grow(nnbr_, u * 8, size_t(cap) * 8);
	grow(nm2l_, u * 27, size_t(cap) * 27);
	grow(np2l_, u * 5, size_t(cap) * 5);
	node_cap_ = cap;
}

static TreeView view(DeviceBuffer<double>& cx, DeviceBuffer<double>& cy, DeviceBuffer<double>& hl, DeviceBuffer<int>& lv,
	DeviceBuffer<int>& par, DeviceBuffer<int>& c0, DeviceBuffer<int>& b, DeviceBuffer<int>& n, DeviceBuffer<int>& nn,
	DeviceBuffer<int>& nm, DeviceBuffer<int>& np, DeviceBuffer<int>& nb, DeviceBuffer<int>& m2l, DeviceBuffer<int>& p2l)
{
	return TreeView{cx.ptr, cy.ptr, hl.ptr, lv.ptr, par.ptr, c0.ptr, b.ptr, n.ptr, nn.ptr, nm.ptr, np.ptr, nb.ptr, m2l.ptr, p2l.ptr};
}

void Engine::build_tree(double root_half_length)
{
	const int n = nstars_;
	num_nodes_ = 0;
	ensure_node_capacity(std::max(1024, n / 4));
	level_start_.assign({0, 1});
	levels_ = 0;

	// root (ipm.cuh:746-773)
	{
		double g[3] = {0.0, 0.0, root_half_length};
		int iv[11] = {0, -1, -1, 0, n, 0, 0, 0};
		MLENS_HIP_CHECK(hipMemcpy(ncx_.ptr, &g[0], sizeof(double), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(ncy_.ptr, &g[1], sizeof(double), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nhl_.ptr, &g[2], sizeof(double), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nlevel_.ptr, &iv[0], sizeof(int), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nparent_.ptr, &iv[1], sizeof(int), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nchild0_.ptr, &iv[2], sizeof(int), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nbegin_.ptr, &iv[3], sizeof(int), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(ncount_.ptr, &iv[4], sizeof(int), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nnnbr_.ptr, &iv[5], sizeof(int), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nnm2l_.ptr, &iv[6], sizeof(int), hipMemcpyHostToDevice));
		MLENS_HIP_CHECK(hipMemcpy(nnp2l_.ptr, &iv[7], sizeof(int), hipMemcpyHostToDevice));
	}
	num_nodes_ = 1;

	star_node_.alloc(std::max(n, 1));
	star_node_tmp_.alloc(std::max(n, 1));
	tx_.alloc(std::max(n, 1));
	ty_.alloc(std::max(n, 1));
	tm_.alloc(std::max(n, 1));
	MLENS_HIP_CHECK(hipMemsetAsync(star_node_.ptr, 0, std::max(n, 1) * sizeof(int), stream_));

	DeviceBuffer<int> maxcnt;
	maxcnt.alloc(1);
	// scratch: flags/rank for nodes, onehot/pre for stars, tiles
	DeviceBuffer<char> sbuf;

	while (true)
	{
		int l0 = level_start_[levels_];
		int l1 = level_start_[levels_ + 1];
		int nl = l1 - l0;
		size_t need = sizeof(I4) * (size_t(nl) + size_t(nl) + 1 + size_t(n) + size_t(n) + 1 +
			2 * (size_t(std::max(n, nl)) / SCAN_TILE + 4));
		sbuf.alloc(need);
		I4* flags = reinterpret_cast<I4*>(sbuf.ptr);
		I4* rank = flags + nl;
		I4* onehot = rank + nl + 1;
		I4* pre = onehot + n;
		I4* tiles = pre + n + 1;

		TreeView t = view(ncx_, ncy_, nhl_, nlevel_, nparent_, nchild0_, nbegin_, ncount_, nnnbr_, nnm2l_, nnp2l_, nnbr_, nm2l_, np2l_);
		MLENS_HIP_CHECK(hipMemsetAsync(maxcnt.ptr, 0, sizeof(int), stream_));
		k_near_count<<<blocks_for(nl, 256), 256, 0, stream_>>>(t, l0, l1, flags, maxcnt.ptr);
		MLENS_HIP_CHECK(hipGetLastError());
		int hmax = 0;
		MLENS_HIP_CHECK(hipMemcpyAsync(&hmax, maxcnt.ptr, sizeof(int), hipMemcpyDeviceToHost, stream_));
		MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
		if (hmax <= MAX_NUM_STARS_DIRECT) break;

		scan_i4(flags, rank, tiles, nl, stream_);
		I4 tot;
		MLENS_HIP_CHECK(hipMemcpyAsync(&tot, rank + nl, sizeof(I4), hipMemcpyDeviceToHost, stream_));
		MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
		int nsplit = tot.x;
		int base = l1;
		ensure_node_capacity(base + 4 * nsplit);
		t = view(ncx_, ncy_, nhl_, nlevel_, nparent_, nchild0_, nbegin_, ncount_, nnnbr_, nnm2l_, nnp2l_, nnbr_, nm2l_, np2l_);

		k_make_children<<<blocks_for(nl, 256), 256, 0, stream_>>>(t, l0, l1, flags, rank, base);
		k_star_onehot<<<blocks_for(n, 256), 256, 0, stream_>>>(t, sx_.ptr, sy_.ptr, star_node_.ptr, l0, flags, onehot, n);
		MLENS_HIP_CHECK(hipGetLastError());
		scan_i4(onehot, pre, tiles, n, stream_);
		k_child_ranges<<<blocks_for(nl, 256), 256, 0, stream_>>>(t, l0, l1, flags, pre);
		k_star_scatter<<<blocks_for(n, 256), 256, 0, stream_>>>(t, sx_.ptr, sy_.ptr, sm_.ptr, tx_.ptr, ty_.ptr, tm_.ptr,
			star_node_.ptr, star_node_tmp_.ptr, l0, flags, pre, n);
		MLENS_HIP_CHECK(hipGetLastError());
		std::swap(sx_.ptr, tx_.ptr);
		std::swap(sy_.ptr, ty_.ptr);
		std::swap(sm_.ptr, tm_.ptr);
		std::swap(star_node_.ptr, star_node_tmp_.ptr);

		num_nodes_ = base + 4 * nsplit;
		levels_++;
		level_start_.push_back(num_nodes_);
		k_set_neighbors<<<blocks_for(4 * nsplit, 256), 256, 0, stream_>>>(t, base, num_nodes_);
		MLENS_HIP_CHECK(hipGetLastError());
	}
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::export_tree(double* geom, int* info, int* lists) const
{
	int nn = num_nodes_;
	std::vector<double> cx(nn), cy(nn), hl(nn);
	std::vector<int> lv(nn), par(nn), c0(nn), b(nn), n(nn), nnb(nn), nm(nn), np(nn), nb(size_t(nn) * 8), m2l(size_t(nn) * 27),
		p2l(size_t(nn) * 5);
	auto get = [](auto& v, const auto& d, size_t cnt) {
		MLENS_HIP_CHECK(hipMemcpy(v.data(), d.ptr, cnt * sizeof(v[0]), hipMemcpyDeviceToHost));
	};
	get(cx, ncx_, nn);
	get(cy, ncy_, nn);
	get(hl, nhl_, nn);
	get(lv, nlevel_, nn);
	get(par, nparent_, nn);
	get(c0, nchild0_, nn);
	get(b, nbegin_, nn);
	get(n, ncount_, nn);
	get(nnb, nnnbr_, nn);
	get(nm, nnm2l_, nn);
	get(np, nnp2l_, nn);
	get(nb, nnbr_, size_t(nn) * 8);
	get(m2l, nm2l_, size_t(nn) * 27);
	get(p2l, np2l_, size_t(nn) * 5);
	for (int i = 0; i < nn; i++)
	{
		geom[3 * i] = cx[i];
		geom[3 * i + 1] = cy[i];
		geom[3 * i + 2] = hl[i];
		int* f = info + 9 * i;
		f[0] = lv[i];
		f[1] = par[i];
		f[2] = c0[i];
		f[3] = c0[i] >= 0 ? 4 : 0;
		f[4] = b[i];
		f[5] = n[i];
		f[6] = nnb[i];
		f[7] = nm[i];
		f[8] = np[i];
		int* L = lists + 40 * i;
		for (int j = 0; j < 40; j++) L[j] = -1;
		for (int j = 0; j < nnb[i]; j++) L[j] = nb[8 * size_t(i) + j];
		for (int j = 0; j < nm[i]; j++) L[8 + j] = m2l[27 * size_t(i) + j];
		for (int j = 0; j < np[i]; j++) L[35 + j] = p2l[5 * size_t(i) + j];
	}
}

/* ---------------------------------------------------------------------------
FMM in node-normalised units (fmm.cuh:21-464). One thread per (node, order).
Binomial coefficients are exact integers rounded once to double; the
reference's int32 table wraps for rows >= 34 (binomial_coefficients.cuh:11-44).
--------------------------------------------------------------------------- */
__device__ __forceinline__ double binom(const double* B, int n, int k)
{
	return (n < k) ? 0.0 : B[n * (n + 1) / 2 + k];
}

__device__ __forceinline__ dcx ld(const double* a, long long i) { return dcx{a[2 * i], a[2 * i + 1]}; }
__device__ __forceinline__ void st(double* a, long long i, dcx v)
{
	a[2 * i] = v.re;
	a[2 * i + 1] = v.im;
}

// P2M at leaves holding stars (fmm.cuh:21-98)
__global__ void k_p2m(TreeView t, int l0, int l1, int p, const double* sx, const double* sy, const double* sm, double* mult)
{
	long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int i = l0 + int(g / (p + 1));
	int k = int(g % (p + 1));
	if (i >= l1) return;
	if (t.child0[i] >= 0 || t.count[i] == 0) return;
	double cx = t.cx[i], cy = t.cy[i], h = t.hl[i];
	int b = t.begin[i], n = t.count[i];
	dcx r = {0, 0};
	if (k == 0)
	{
		for (int s = b; s < b + n; s++) r.re += sm[s];
	}
	else
	{
		for (int s = b; s < b + n; s++)
		{
			dcx dz = dcx{(sx[s] - cx) / h, (sy[s] - cy) / h};
			dcx pw = dev::cpow(dz, k);
			r = dev::sub(r, dev::scale(pw, sm[s]));
		}
		r = dev::scale(r, 1.0 / k);
	}
	st(mult, (long long)i * (p + 1) + k, r);
}

// M2M, children to parent (fmm.cuh:109-192)
__global__ void k_m2m(TreeView t, int l0, int l1, int p, const double* B, double* mult)
{
	long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int i = l0 + int(g / (p + 1));
	int l = int(g % (p + 1));
	if (i >= l1 || t.child0[i] < 0) return;
	dcx acc = {0, 0};
	for (int c = t.child0[i]; c < t.child0[i] + 4; c++)
	{
		const long long cb = (long long)c * (p + 1);
		if (l == 0)
		{
			acc = dev::add(acc, ld(mult, cb));
			continue;
		}
		dcx dz = dcx{(t.cx[c] - t.cx[i]) / t.hl[c], (t.cy[c] - t.cy[i]) / t.hl[c]};
		dcx idz = dev::cdiv(dcx{1, 0}, dz);
		dcx r = {0, 0};
		for (int k = l; k >= 1; k--)
		{
			r = dev::add(r, dev::scale(ld(mult, cb + k), binom(B, l - 1, k - 1)));
			r = dev::mul(r, idz);
		}
		r = dev::sub(r, dev::scale(ld(mult, cb), 1.0 / l));
		r = dev::mul(r, dev::cpow(dz, l));
		r = dev::scale(r, ldexp(1.0, -l));
		acc = dev::add(acc, r);
	}
	st(mult, (long long)i * (p + 1) + l, acc);
}

// L2L (overwrite), M2L and P2L for one level (fmm.cuh:204-464)
__global__ void k_local(TreeView t, int l0, int l1, int p, const double* B, const double* mult, double* local,
	const double* sx, const double* sy, const double* sm, int do_m2l)
{
	long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int i = l0 + int(g / (p + 1));
	int l = int(g % (p + 1));
	if (i >= l1) return;
	const double cx = t.cx[i], cy = t.cy[i], h = t.hl[i];
	dcx acc;
	{
		int par = t.parent[i];
		const long long pb = (long long)par * (p + 1);
		dcx dz = dcx{(cx - t.cx[par]) / t.hl[par], (cy - t.cy[par]) / t.hl[par]};
		dcx r = {0, 0};
		for (int k = p; k >= 1; k--)
		{
			r = dev::add(r, dev::scale(ld(local, pb + k), binom(B, k, l)));
			r = dev::mul(r, dz);
		}
		r = dev::cdiv(r, dev::cpow(dz, l));
		acc = dev::scale(r, ldexp(1.0, -l));
	}
	for (int j = 0; j < (do_m2l ? t.nm2l[i] : 0); j++)
	{
		int f = t.m2l[27 * i + j];
		const long long fb = (long long)f * (p + 1);
		double hf = t.hl[f];
		dcx dz = dcx{(t.cx[f] - cx) / hf, (t.cy[f] - cy) / hf};
		dcx mdz = dcx{-dz.re, -dz.im};
		dcx imdz = dev::cdiv(dcx{1, 0}, mdz);
		dcx b0 = ld(mult, fb);
		dcx r = {0, 0};
		if (l == 0)
		{
			for (int k = p; k >= 1; k--) r = dev::mul(dev::add(r, ld(mult, fb + k)), imdz);
			r = dev::add(r, dev::mul(b0, dev::clog(mdz)));
			r = dev::add(r, dev::scale(b0, log(hf)));
		}
		else
		{
			for (int k = p; k >= 1; k--) r = dev::mul(dev::add(r, dev::scale(ld(mult, fb + k), binom(B, l + k - 1, k - 1))), imdz);
			r = dev::sub(r, dev::scale(b0, 1.0 / l));
			r = dev::cdiv(r, dev::cpow(dz, l));
		}
		acc = dev::add(acc, r);
	}
	for (int j = 0; j < t.np2l[i]; j++)
	{
		int f = t.p2l[5 * i + j];
		for (int s = t.begin[f]; s < t.begin[f] + t.count[f]; s++)
		{
			dcx dz = dcx{(sx[s] - cx) / h, (sy[s] - cy) / h};
			double m = sm[s];
			if (l == 0)
			{
				acc = dev::add(acc, dev::scale(dev::clog(dcx{-dz.re, -dz.im}), m));
				acc.re += m * log(h);
			}
			else
			{
				acc = dev::sub(acc, dev::cdiv(dcx{m, 0}, dev::scale(dev::cpow(dz, l), double(l))));
			}
		}
	}
	st(local, (long long)i * (p + 1) + l, acc);
}

// leaf flags for compaction
__global__ void k_leaf_flags(const int* child0, int nn, I4* flags)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < nn) flags[i] = I4{child0[i] < 0 ? 1 : 0, 0, 0, 0};
}

/*
Leaf tables for shooting: geometry, the near-field star list (own stars then
each neighbour's, alpha_star.cuh:26-37) and the deflection polynomial
A_i = (theta^2/h) i c_i of alpha_local.cuh:18-36, so that
alpha_local(z) = conj( sum_{i=1..p} A_i dz^(i-1) ), dz = (z - c)/h.
*/
__global__ void k_leaf_tables(TreeView t, int nn, const I4* pre, int p, double theta2, const double* sx, const double* sy,
	const double* sm, const double* local, int* leaf_of_node, double* geom, int* near_n, int* near_true, double* near_d,
	float* near_f, double* poly_d, float* poly_f)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nn) return;
	if (t.child0[i] >= 0)
	{
		leaf_of_node[i] = -1;
		return;
	}
	int L = pre[i].x;
	leaf_of_node[i] = L;
	double cx = t.cx[i], cy = t.cy[i], h = t.hl[i];
	geom[4 * (long long)L] = cx;
	geom[4 * (long long)L + 1] = cy;
	geom[4 * (long long)L + 2] = h;
	geom[4 * (long long)L + 3] = 1.0 / h;
	int k = 0;
	auto put = [&](int b, int n) {
		for (int s = b; s < b + n && k < MAX_NUM_STARS_DIRECT; s++, k++)
		{
			long long o = ((long long)L * MAX_NUM_STARS_DIRECT + k) * 3;
			near_d[o] = sx[s];
			near_d[o + 1] = sy[s];
			near_d[o + 2] = sm[s];
			long long of = ((long long)L * MAX_NUM_STARS_DIRECT + k) * 4;
			near_f[of] = float(sx[s] - cx);
			near_f[of + 1] = float(sy[s] - cy);
			near_f[of + 2] = float(sm[s]);
			near_f[of + 3] = 0.0f;
		}
	};
	put(t.begin[i], t.count[i]);
	for (int j = 0; j < t.nnbr[i]; j++)
	{
		int nb = t.nbr[8 * i + j];
		put(t.begin[nb], t.count[nb]);
	}
	near_true[L] = k;
	// pad to a multiple of 4 with massless stars far away (their terms are exactly 0)
	for (; (k & 3) != 0; k++)
	{
		long long o = ((long long)L * MAX_NUM_STARS_DIRECT + k) * 3;
		near_d[o] = 1e150;
		near_d[o + 1] = 1e150;
		near_d[o + 2] = 0.0;
		long long of = ((long long)L * MAX_NUM_STARS_DIRECT + k) * 4;
		near_f[of] = 1e18f;
		near_f[of + 1] = 1e18f;
		near_f[of + 2] = 0.0f;
		near_f[of + 3] = 0.0f;
	}
	// polynomial tables in descending degree, zero-padded at the top to a
	// multiple of 4: entry j of the FP64 table is the coefficient of
	// dz^(SD-1-j) in P(dz) = sum_{i=1..p} A_i dz^(i-1); entry j of the FP32
	// table is the coefficient of dz^(SF-1-j) in Q(dz) = sum_{i=2..p} A_i dz^(i-2)
	const int SD = poly_stride_d(p), SF = poly_stride_f(p);
	static_assert((MAX_EXPANSION_ORDER - 1 + 3) / 4 * 4 <= 32, "FP32 polynomial longer than its bound table");
	const long long lb = (long long)i * (p + 1);
	for (int j = 0; j < SD; j++)
	{
		int q = SD - j;   // A_q
		dcx a = (q <= p) ? dev::scale(ld(local, lb + q), theta2 / h * q) : dcx{0, 0};
		poly_d[2 * ((long long)L * SD + j)] = a.re;
		poly_d[2 * ((long long)L * SD + j) + 1] = a.im;
	}
	// the FP32 evaluation skips its leading groups of 4 entries while their
	// terms together stay below 2^-28 of the scale of all terms, bounded over
	// the leaf (|dz| <= sqrt(2) in units of the half-size): an eighth of one
	// rounding of the FP32 sum, which carries about SF such roundings already.
	// On the 10000^2 map (p = 30) this leaves the mixed map's error against FP64
	// as it was (median 2.1e-6, max 6.9e-4) and takes 1% off the shot; at 2^-25
	// single pixels moved by 3e-3.
	const double R = 1.4143;
	double S = dev::cabs(dev::scale(ld(local, lb + 1), theta2 / h)), lead[32];   // SF <= 32 up to MAX_EXPANSION_ORDER
	for (int j = 0; j < SF; j++)
	{
		int q = SF - j + 1;   // A_q
		dcx a = (q <= p && q >= 2) ? dev::scale(ld(local, lb + q), theta2 / h * q) : dcx{0, 0};
		poly_f[2 * ((long long)L * SF + j)] = float(a.re);
		poly_f[2 * ((long long)L * SF + j) + 1] = float(a.im);
		lead[j] = dev::cabs(a) * pow(R, double(q - 1));   // bound of A_q dz^(q-1) over the leaf
		S += lead[j];
	}
	int skip = 0;
	double tail = 0;
	for (int j0 = 0; j0 + 4 <= SF; j0 += 4)
	{
		tail += lead[j0] + lead[j0 + 1] + lead[j0 + 2] + lead[j0 + 3];
		if (!(tail <= 0x1p-28 * S)) break;
		skip = j0 + 4;
	}
	near_n[L] = k | (skip << 16);
}

void Engine::fmm(int p, const LensParams& lens)
{
	p_ = p;
	// exact Pascal triangle rows 0..2p (C(62, 31) < 2^64)
	{
		int nmax = 2 * p;
		std::vector<double> B(size_t(nmax + 1) * (nmax + 2) / 2);
		std::vector<unsigned long long> row(nmax + 1), prev(nmax + 1);
		for (int nr = 0; nr <= nmax; nr++)
		{
			for (int k = 0; k <= nr; k++) row[k] = (k == 0 || k == nr) ? 1ull : prev[k - 1] + prev[k];
			for (int k = 0; k <= nr; k++) B[size_t(nr) * (nr + 1) / 2 + k] = double(row[k]);
			prev = row;
		}
		binom_.alloc(B.size());
		MLENS_HIP_CHECK(hipMemcpy(binom_.ptr, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice));
	}
	const size_t nc = size_t(num_nodes_) * (p + 1) * 2;
	mult_.alloc(nc);
	local_.alloc(nc);
	MLENS_HIP_CHECK(hipMemsetAsync(mult_.ptr, 0, nc * sizeof(double), stream_));
	MLENS_HIP_CHECK(hipMemsetAsync(local_.ptr, 0, nc * sizeof(double), stream_));
	TreeView t = view(ncx_, ncy_, nhl_, nlevel_, nparent_, nchild0_, nbegin_, ncount_, nnnbr_, nnm2l_, nnp2l_, nnbr_, nm2l_, np2l_);
	for (int lev = levels_; lev >= 0; lev--)
	{
		int l0 = level_start_[lev], l1 = level_start_[lev + 1];
		long long work = (long long)(l1 - l0) * (p + 1);
		k_p2m<<<blocks_for(work, 256), 256, 0, stream_>>>(t, l0, l1, p, sx_.ptr, sy_.ptr, sm_.ptr, mult_.ptr);
		k_m2m<<<blocks_for(work, 256), 256, 0, stream_>>>(t, l0, l1, p, binom_.ptr, mult_.ptr);
		MLENS_HIP_CHECK(hipGetLastError());
	}
	// M2L: FP64 recurrences inside k_local, or batched GEMMs on MFMA after it
	const bool mfma = m2l_mode_ != M2L_DIRECT;
	const bool f32 = m2l_mode_ == M2L_MFMA_F32 || (m2l_mode_ == M2L_MFMA && p <= 20);
	if (mfma)
	{
		m2l_prepare(p, f32);
		if (m2l_count_.count < 2) m2l_count_.alloc(2);
		MLENS_HIP_CHECK(hipMemsetAsync(m2l_count_.ptr, 0, 2 * sizeof(unsigned long long), stream_));
		while (m2l_ev_.size() < size_t(2 * (levels_ + 1)))
		{
			hipEvent_t e;
			MLENS_HIP_CHECK(hipEventCreate(&e));
			m2l_ev_.push_back(e);
		}
	}
	for (int lev = 2; lev <= levels_; lev++)
	{
		int l0 = level_start_[lev], l1 = level_start_[lev + 1];
		long long work = (long long)(l1 - l0) * (p + 1);
		k_local<<<blocks_for(work, 256), 256, 0, stream_>>>(t, l0, l1, p, binom_.ptr, mult_.ptr, local_.ptr, sx_.ptr,
			sy_.ptr, sm_.ptr, mfma ? 0 : 1);
		MLENS_HIP_CHECK(hipGetLastError());
		if (mfma) m2l_level(lev, p, f32, m2l_ev_[2 * lev], m2l_ev_[2 * lev + 1]);
	}
	build_leaf_tables(lens);
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	m2l_ms_ = 0;
	m2l_pairs_ = 0;
	m2l_f32_ = mfma && f32;
	if (mfma)
	{
		for (int lev = 2; lev <= levels_; lev++)
		{
			float ms = 0;
			MLENS_HIP_CHECK(hipEventElapsedTime(&ms, m2l_ev_[2 * lev], m2l_ev_[2 * lev + 1]));
			m2l_ms_ += ms;
		}
		unsigned long long h[2];
		MLENS_HIP_CHECK(hipMemcpy(h, m2l_count_.ptr, sizeof(h), hipMemcpyDeviceToHost));
		m2l_pairs_ = (long long)h[0];
		if (h[1] & 1) throw HipError("an M2L source is not at one of the 40 offsets of its target's level", __FILE__, __LINE__);
	}
}

void Engine::build_leaf_tables(const LensParams& lens)
{
	const int nn = num_nodes_;
	DeviceBuffer<char> sbuf;
	sbuf.alloc(sizeof(I4) * (2 * size_t(nn) + 1 + 2 * (size_t(nn) / SCAN_TILE + 4)));
	I4* flags = reinterpret_cast<I4*>(sbuf.ptr);
	I4* pre = flags + nn;
	I4* tiles = pre + nn + 1;
	k_leaf_flags<<<blocks_for(nn, 256), 256, 0, stream_>>>(nchild0_.ptr, nn, flags);
	scan_i4(flags, pre, tiles, nn, stream_);
	I4 tot;
	MLENS_HIP_CHECK(hipMemcpyAsync(&tot, pre + nn, sizeof(I4), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	num_leaves_ = tot.x;
	const int SD = poly_stride_d(p_), SF = poly_stride_f(p_);
	leaf_of_node_.alloc(nn);
	leaf_geom_.alloc(size_t(num_leaves_) * 4);
	leaf_near_n_.alloc(num_leaves_);
	leaf_near_true_.alloc(num_leaves_);
	leaf_near_d_.alloc(size_t(num_leaves_) * MAX_NUM_STARS_DIRECT * 3);
	leaf_near_f_.alloc(size_t(num_leaves_) * MAX_NUM_STARS_DIRECT * 4);
	leaf_poly_d_.alloc(size_t(std::max(num_leaves_, 1)) * SD * 2);
	leaf_poly_f_.alloc(size_t(std::max(num_leaves_, 1)) * std::max(SF, 1) * 2);
	TreeView t = view(ncx_, ncy_, nhl_, nlevel_, nparent_, nchild0_, nbegin_, ncount_, nnnbr_, nnm2l_, nnp2l_, nnbr_, nm2l_, np2l_);
	k_leaf_tables<<<blocks_for(nn, 256), 256, 0, stream_>>>(t, nn, pre, p_, lens.theta * lens.theta, sx_.ptr, sy_.ptr,
		sm_.ptr, local_.ptr, leaf_of_node_.ptr, leaf_geom_.ptr, leaf_near_n_.ptr, leaf_near_true_.ptr, leaf_near_d_.ptr,
		leaf_near_f_.ptr,
		leaf_poly_d_.ptr, leaf_poly_f_.ptr);
	MLENS_HIP_CHECK(hipGetLastError());
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::export_coeffs(int which, double* out) const
{
	const DeviceBuffer<double>& b = which ? local_ : mult_;
	MLENS_HIP_CHECK(hipMemcpy(out, b.ptr, size_t(num_nodes_) * (p_ + 1) * 2 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace mlens
