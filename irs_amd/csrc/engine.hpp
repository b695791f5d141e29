// Device engine: star field, quadtree, FMM and the shooting kernels, all in HBM.
//
// Layout (everything structure-of-arrays, FP64 unless noted):
//   stars      sx, sy, sm [N], permuted into tree order by build_tree
//   nodes      flat arrays over all levels; level l occupies
//              [level_start[l], level_start[l+1]); children of a node are four
//              consecutive nodes starting at child0 (tree_node.cuh:151-177)
//   lists      nbr[8], m2l[27], p2l[5] per node (tree_node.cuh:41-49)
//   coeffs     mult, local: (p+1) complex per node (tree_node.cuh:54-56)
//   leaves     compact tables read by the shooting kernels: centre, 1/h, the
//              near-field star list (own + neighbours, <= 32, stride 32) and
//              the deflection polynomial A_i = (theta^2/h) i c_i, i = 1..p
#pragma once

#include "common.hpp"

#include <vector>

namespace mlens
{

// strides (in complex entries) of the per-leaf deflection polynomials
__host__ __device__ inline int poly_stride_d(int p) { return (p + 3) / 4 * 4; }
__host__ __device__ inline int poly_stride_f(int p) { return p >= 2 ? (p - 1 + 3) / 4 * 4 : 0; }

enum Precision
{
	PREC_FP64 = 0,    // every stage in FP64
	PREC_MIXED = 1,   // FP64 anchors, FP32 leaf-relative near field / Horner / pixel-local polygons
};

// How the M2L translations of the FMM are evaluated
enum M2LMode
{
	M2L_DIRECT = 0,     // FP64 recurrences per (node, order) thread (fmm.cuh:281-374 order of operations)
	M2L_MFMA = 1,       // batched GEMMs on MFMA: FP32 for p <= 20, FP64 above
	M2L_MFMA_F32 = 2,   // batched GEMMs on FP32 MFMA at any p
	M2L_MFMA_F64 = 3,   // batched GEMMs on FP64 MFMA at any p
};

enum ShootKernel
{
	KERNEL_FAST = 0,        // shared-corner lattice tiles, wave-uniform leaf tables
	KERNEL_REFERENCE = 1,   // the reference's per-cell algorithm, FP64 (ipm_functions.cuh:112-230)
};

constexpr int BATCH_HIST_BINS = 128;   // log10 magnification histogram of a batched map, [-2, 2]

template <typename T>
struct DeviceBuffer
{
	T* ptr = nullptr;
	size_t count = 0;
	void alloc(size_t n);
	void free();
	~DeviceBuffer() { free(); }
	DeviceBuffer() = default;
	DeviceBuffer(const DeviceBuffer&) = delete;
	DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

struct StarStats
{
	double m_tot, m2_tot, m2_ln_tot, m_min, m_max;
};

struct CCFResult
{
	double t_init_ms = 0, t_ccs_ms = 0, t_caustics_ms = 0;   // starting roots, phase sweep, outputs (HIP events)
	double max_error = 0;   // largest error bound on 1/mu over all roots
	bool has_nan = false;   // some error bound is not finite
};

struct ShootStats
{
	double ms = 0;                 // kernel time (HIP events)
	long long cells = 0;           // cells (IPM) or rays (IRS) shot
	long long rays_received = 0;   // IRS only
};

class Engine
{
public:
	Engine();
	~Engine();
	Engine(const Engine&) = delete;
	Engine& operator=(const Engine&) = delete;

	hipStream_t stream() const { return stream_; }

	/* stars */
	void upload_stars(const double* aos, int n);
	void generate_stars(int n, int rectangular, double c_re, double c_im, int mass_function, double m_lower,
		double m_upper, double m_solar, uint32_t seed);
	void download_stars(double* aos) const;   // AoS x1, x2, m in the current (tree) order
	StarStats star_stats();
	int num_stars() const { return nstars_; }

	/* quadtree with the reference's subdivision and interaction-list rules */
	void build_tree(double root_half_length);
	int tree_levels() const { return levels_; }
	int num_nodes() const { return num_nodes_; }
	const std::vector<int>& level_start() const { return level_start_; }
	void export_tree(double* geom, int* info, int* lists) const;   // same layout as the oracle's export

	/* FMM (P2M, M2M, L2L, M2L, P2L) and the leaf tables for shooting */
	void fmm(int p, const LensParams& lens);
	int expansion_order() const { return p_; }
	void set_m2l_mode(int mode) { m2l_mode_ = mode; }
	int m2l_mode() const { return m2l_mode_; }
	double m2l_ms() const { return m2l_ms_; }              // GPU time of the last fmm()'s M2L kernels
	long long m2l_pairs() const { return m2l_pairs_; }     // M2L (target, source) pairs of the last fmm()
	bool m2l_used_f32() const { return m2l_f32_; }
	void export_coeffs(int which, double* out) const;   // 0 = multipole, 1 = local
	double mean_near_stars(const GridParams& g) const;  // lattice-weighted mean near-field count

	/* maps (row-major n2 x n1, row 0 at the top) */
	void alloc_maps(int n1, int n2, bool ipm, Precision prec, bool parities);
	void zero_maps();
	ShootStats shoot_ipm(const LensParams& lens, const GridParams& g, Precision prec, ShootKernel kern,
		long long row0 = 0, long long row1 = -1);
	ShootStats shoot_irs(const LensParams& lens, const GridParams& g, Precision prec, long long row0 = 0,
		long long row1 = -1);
	/* lattice rows per tile of the shooting kernels: 0 picks them per launch
	   (IPM 64-1024, IRS 32-256); a positive value forces that height */
	void set_tile_rows(int rows) { tile_rows_ = rows; }
	int tile_rows() const { return tile_rows_; }
	int last_tile_rows() const { return last_tile_rows_; }   // the height of the last shooting launch
	void add_parity_maps();
	void download_map(int which, void* host) const;   // 0 = total, 1 = minima, 2 = saddles
	void map_minmax(int which, double* mn, double* mx);
	void map_histogram(int which, bool log10_of_value, int factor, long long hist_min, long long length,
		long long* host_hist);

	/* critical curves (ccf_functions.cuh, ccf.cuh): refine nroots roots per phase for
	   nphi + nb phases (nb branches of nphi / nb + 1 phases), starting from init_roots
	   (nb x nroots complex, the roots at each branch's centre phase). Outputs are
	   root-major, nroots x (nphi + nb): ccs and caustics complex (re, im), mu_scales
	   real; caustics and mu_scales may be null. Needs build_tree and fmm first. */
	CCFResult ccf_solve(const LensParams& lens, const double* init_roots, int nroots, int nphi, int nb, bool circle_pole,
		double* ccs, double* caustics, double* mu_scales, int init_iters = 30, int step_iters = 20);

	/* thin test entry: lens equation at arbitrary points through the tree */
	void deflect_points(const LensParams& lens, const double* z, int n, Precision prec, double* w, double* mu);
	/* the same without allocating: the points, deflections and magnifications
	   are staged in a fixed device array of DEFLECT_NOALLOC_MAX points (one per
	   device, kept until the process ends); calls that overlap on one device race */
	static constexpr int DEFLECT_NOALLOC_MAX = 1 << 16;
	void deflect_points_noalloc(const LensParams& lens, const double* z, int n, Precision prec, double* w, double* mu);

	/* the per-ray source position of IRS rows (test entry for bit-exact binning) */
	void irs_positions(const LensParams& lens, const GridParams& g, Precision prec, long long row, double* w);

	int map_pixels(int i) const { return npix_[i]; }
	bool maps_are_fp32() const { return map_fp32_; }

	/* lattice cells whose mixed-precision corners' pixel box meets columns [wx0, wx0 + ww) and rows
	   [wy0, wy0 + wh) widened by margin pixels: records (lattice row, first column of a 63-cell tile,
	   mask of its cells), at most cap written to rec (3 x cap values); returns the number found */
	long long window_cells(const LensParams& lens, const GridParams& g, int wx0, int wy0, int ww, int wh, double margin,
		long long* rec, long long cap);
	/* kernel arguments of this engine's IPM map (mixed precision) for a batched launch;
	   `out` receives the engine-internal argument block, batch_args_size() bytes */
	static size_t batch_args_size();
	void fill_batch_args(void* out, const LensParams& lens, const GridParams& g) const;

private:
	void ensure_node_capacity(int n);
	void build_leaf_tables(const LensParams& lens);
	void m2l_prepare(int p, bool f32);
	void m2l_level(int lev, int p, bool f32, hipEvent_t e0, hipEvent_t e1);

	hipStream_t stream_ = nullptr;
	int num_cus_ = 0;   // compute units of the current device
	int dev_ = 0;       // the device current when the engine was made
	int tile_rows_ = 0, last_tile_rows_ = 0;
	hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

	int nstars_ = 0;
	DeviceBuffer<double> sx_, sy_, sm_, tx_, ty_, tm_;
	DeviceBuffer<int> star_node_, star_node_tmp_;

	int levels_ = 0;
	int num_nodes_ = 0;
	int node_cap_ = 0;
	std::vector<int> level_start_;
	DeviceBuffer<double> ncx_, ncy_, nhl_;
	DeviceBuffer<int> nlevel_, nparent_, nchild0_, nbegin_, ncount_, nnnbr_, nnm2l_, nnp2l_;
	DeviceBuffer<int> nnbr_, nm2l_, np2l_;

	int p_ = 0;
	DeviceBuffer<double> binom_;
	DeviceBuffer<double> mult_, local_;   // interleaved re, im

	// MFMA M2L: the 40 offset matrices (real-stacked, transposed, padded),
	// per-level source tables and real-stacked multipoles
	int m2l_mode_ = M2L_DIRECT;
	double m2l_ms_ = 0;
	long long m2l_pairs_ = 0;
	bool m2l_f32_ = false;
	int m2l_ready_p_ = -1;
	bool m2l_ready_f32_ = false;
	DeviceBuffer<float> m2l_af_;
	DeviceBuffer<double> m2l_ad_;
	DeviceBuffer<int> m2l_src_;
	DeviceBuffer<float> mstack_f_;
	DeviceBuffer<double> mstack_d_;
	DeviceBuffer<unsigned long long> m2l_count_;
	std::vector<hipEvent_t> m2l_ev_;

	int num_leaves_ = 0;
	DeviceBuffer<int> leaf_of_node_;
	DeviceBuffer<double> leaf_geom_;      // cx, cy, hl, 1/hl per leaf
	DeviceBuffer<int> leaf_near_n_;       // padded to a multiple of 4; FP32 polynomial entries to skip << 16
	DeviceBuffer<int> leaf_near_true_;    // stars actually in the list
	DeviceBuffer<double> leaf_near_d_;    // x, y, m (FP64), stride 32 per leaf
	DeviceBuffer<float> leaf_near_f_;     // x - cx, y - cy, m (FP32), stride 32 per leaf
	DeviceBuffer<double> leaf_poly_d_;    // A_p..A_1 (re, im), descending, zero-padded to a multiple of 4
	DeviceBuffer<float> leaf_poly_f_;     // A_p..A_2 (re, im), descending, zero-padded to a multiple of 4

	int npix_[2] = {0, 0};
	bool map_fp32_ = false;
	bool map_int_ = false;
	bool parities_ = false;
	DeviceBuffer<char> map_, map_min_, map_sad_;
	// tiled accumulation maps of the fast IPM kernel (all zero between shots)
	DeviceBuffer<char> tmap_, tmap_min_, tmap_sad_;
	// FP64 span maps beside them: span_ for the map, span_min_ / span_sad_ with parities
	DeviceBuffer<char> span_, span_min_, span_sad_;
	int tiles_x_ = 0;
	size_t tmap_bytes_ = 0;
	void ensure_tiled_maps();
	void fold_tiled_maps();

	DeviceBuffer<int> scratch_i_;
	DeviceBuffer<double> scratch_d_;
	DeviceBuffer<long long> scratch_ll_;
};

/*
Many IPM maps shot in one persistent launch. Each instance is an Engine whose
tree and FMM are built and whose FP32 map is allocated; the batch zeroes the
maps, shoots every instance's lattice tiles from one work queue and bins each
map into a 128-bin log10 magnification histogram over [-2, 2], all captured
in one HIP graph. The first shoot() times candidate tile heights and resident
blocks per CU on a sample of the tiles and keeps the fastest.
*/
class BatchShooter
{
public:
	BatchShooter();
	~BatchShooter();
	BatchShooter(const BatchShooter&) = delete;
	BatchShooter& operator=(const BatchShooter&) = delete;

	int add(Engine* e, const LensParams& lens, const GridParams& g);   // returns the instance index
	double shoot(int verbose);                                          // ms of the graph launch
	void histograms(unsigned* host) const;                              // size() x BATCH_HIST_BINS
	int size() const;
	long long cells() const;
	int tile_h() const;
	int blocks_per_cu() const;
	double tune_ms() const;
	void set_tuning(int tile_h, int blocks_per_cu);   // skip the autotuner

private:
	struct Impl;
	Impl* impl_;
};

}  // namespace mlens
