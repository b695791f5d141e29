// Host orchestration of one magnification map: the IPM<T> / IRS<T> classes of
// the reference (ipm.cuh:34-1357, irs.cuh), with the same public fields and
// defaults, the same validation and derived-parameter formulas, and run() /
// save(). Differences:
//   * validation and parameter derivation run before any device call, so a
//     bad configuration fails without a GPU (the reference resets the device
//     first, ipm.cuh:1329-1331);
//   * no device reset: other HIP users in the process keep their memory;
//   * arrays returned to callers are host copies that stay valid until this
//     object's next run();
//   * the shooting phase is timed with HIP events, in seconds.
#pragma once

#include "common.hpp"
#include "engine.hpp"
#include "io.hpp"
#include "mass_functions.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <numbers>
#include <string>
#include <vector>

namespace mlens
{

template <typename T, bool IS_IPM>
class MapMaker
{
public:
	/* public input fields and defaults (ipm.cuh:42-64; irs.cuh:50,58 differ) */
	T kappa_tot = static_cast<T>(0.3);
	T shear = static_cast<T>(0.3);
	T kappa_star = static_cast<T>(0.27);
	T theta_star = static_cast<T>(1);
	std::string mass_function_str = "equal";
	T m_solar = static_cast<T>(1);
	T m_lower = static_cast<T>(0.01);
	T m_upper = static_cast<T>(50);
	T light_loss = IS_IPM ? static_cast<T>(0.001) : static_cast<T>(0.01);
	int rectangular = 0;
	int approx = 1;
	T safety_scale = static_cast<T>(1.37);
	std::string starfile = "";
	Complex<T> center_y = Complex<T>(0, 0);
	Complex<T> half_length_y = Complex<T>(5, 5);
	Complex<int> num_pixels_y = Complex<int>(1000, 1000);
	int num_rays_y = IS_IPM ? 1 : 100;
	int random_seed = 0;
	int write_stars = 1;
	int write_maps = 1;
	int write_parities = 0;
	int write_histograms = 1;
	std::string outfile_prefix = "./";

	/* extensions (no reference counterpart) */
	int expansion_order_override = 0;   // 0: automatic order of ipm.cuh:848-860
	int m2l_mode = M2L_MFMA_F64;        // M2LMode: batched GEMMs on FP64 MFMA unless set otherwise
	int kernel = KERNEL_FAST;
	int precision = std::is_same<T, float>::value ? PREC_MIXED : PREC_FP64;
	int tile_rows = 0;                  // lattice rows per tile of the shooting kernel; 0: chosen per launch

	MapMaker() {}

	/* ---------------- staged pipeline ---------------- */
	bool prepare(int verbose)
	{
		prepared_ = false;
		built_ = false;
		try
		{
			if (!check_input_params(verbose)) return false;
			if (!calculate_derived_params(verbose)) return false;
			if (!set_device(verbose)) return false;
			if (!populate_star_array(verbose)) return false;
			stars_stale_ = true;   // the new field, in drawing order until build() sorts it
			prepared_ = true;
			return true;
		}
		catch (const HipError& e)
		{
			std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";
			return false;
		}
	}

	// validation and derived geometry only, nothing on the device (num_stars,
	// lattice and the image-plane region become available)
	bool derive(int verbose)
	{
		prepared_ = false;
		built_ = false;
		return check_input_params(verbose) && calculate_derived_params(verbose);
	}

	// allocates the FP32 map a batched launch writes (needs build() and mixed precision)
	bool prepare_batch_map()
	{
		if (!built_)
		{
			std::cerr << "Error. build() must succeed before the map is batched\n";
			return false;
		}
		if (!IS_IPM || precision != PREC_MIXED || write_parities)
		{
			std::cerr << "Error. batched maps are IPM maps in mixed precision (precision 1) without parities\n";
			return false;
		}
		engine_->alloc_maps(num_pixels_y.re, num_pixels_y.im, true, PREC_MIXED, false);
		return true;
	}

	bool build(int verbose)
	{
		try
		{
			if (!prepared_)
			{
				std::cerr << "Error. prepare() must succeed before build()\n";
				return false;
			}
			built_ = create_tree(verbose);
			stars_stale_ = built_;   // tree order is final once the tree exists; get_stars fetches it
			return built_;
		}
		catch (const HipError& e)
		{
			std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";
			return false;
		}
	}

	bool shoot(int verbose, long long row0 = 0, long long row1 = -1)
	{
		try
		{
			if (!built_)
			{
				std::cerr << "Error. build() must succeed before shoot()\n";
				return false;
			}
			if (!check_input_params(0)) return false;
			// (re)allocates when precision changed since prepare(); zeroes the maps
			engine_->alloc_maps(num_pixels_y.re, num_pixels_y.im, IS_IPM, Precision(precision), write_parities != 0);
			engine_->set_tile_rows(tile_rows);
			LensParams L = lens();
			GridParams G = grid();
			MLENS_VERBOSE((IS_IPM ? "Shooting cells...\n" : "Shooting rays...\n"), verbose, 1);
			ShootStats st;
			if (IS_IPM)
				st = engine_->shoot_ipm(L, G, Precision(precision), ShootKernel(kernel), row0, row1);
			else
				st = engine_->shoot_irs(L, G, Precision(precision), row0, row1);
			t_shoot_ = st.ms / 1000.0;
			cells_shot_ = st.cells;
			num_rays_received_ = st.rays_received;
			MLENS_VERBOSE((IS_IPM ? "Done shooting cells. Elapsed time: " : "Done shooting rays. Elapsed time: ") << t_shoot_
							  << " seconds.\n\n",
				verbose, 1);
			engine_->add_parity_maps();
			return true;
		}
		catch (const HipError& e)
		{
			std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";
			return false;
		}
	}

	bool finish(int verbose)
	{
		try
		{
			if (!built_)
			{
				std::cerr << "Error. shoot() must succeed before finish()\n";
				return false;
			}
			if (!create_histograms(verbose)) return false;
			download_results();
			return true;
		}
		catch (const HipError& e)
		{
			std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";
			return false;
		}
	}

	bool run(int verbose)
	{
		if (!prepare(verbose)) return false;
		if (!build(verbose)) return false;
		if (!shoot(verbose)) return false;
		if (!finish(verbose)) return false;
		return true;
	}

	bool save(int verbose) { return write_files(verbose); }

	/* ---------------- getters ---------------- */
	int get_num_stars() const { return num_stars_; }
	Complex<T> get_corner() const { return rectangular ? corner_ : Complex<T>(corner_.abs(), 0); }   // ipm.cuh:1350
	T* get_stars()
	{
		if (stars_stale_ && engine_)
		{
			try
			{
				download_stars();
			}
			catch (const HipError& e)
			{
				std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";
				return nullptr;
			}
		}
		return stars_host_.empty() ? nullptr : stars_host_.data();
	}
	T* get_pixels() { return pix_.empty() ? nullptr : pix_.data(); }
	T* get_pixels_minima() { return pix_min_.empty() ? nullptr : pix_min_.data(); }
	T* get_pixels_saddles() { return pix_sad_.empty() ? nullptr : pix_sad_.data(); }
	int* get_ipixels() { return ipix_.empty() ? nullptr : ipix_.data(); }
	int* get_ipixels_minima() { return ipix_min_.empty() ? nullptr : ipix_min_.data(); }
	int* get_ipixels_saddles() { return ipix_sad_.empty() ? nullptr : ipix_sad_.data(); }
	double get_t_shoot() const { return t_shoot_; }
	long long get_cells_shot() const { return cells_shot_; }
	long long get_num_rays_received() const { return num_rays_received_; }
	long long get_num_rays_shot() const { return (long long)num_ray_threads_.re * num_ray_threads_.im; }
	int get_tree_levels() const { return tree_levels_; }
	int get_expansion_order() const { return expansion_order_; }
	double get_t_m2l() const { return engine_ ? engine_->m2l_ms() / 1000.0 : 0.0; }
	long long get_m2l_pairs() const { return engine_ ? engine_->m2l_pairs() : 0; }
	int get_m2l_fp32() const { return engine_ && engine_->m2l_used_f32() ? 1 : 0; }
	T get_alpha_error() const { return alpha_error_; }
	T get_mu_ave() const { return mu_ave_; }
	T get_root_half_length() const { return root_half_length_; }
	Complex<T> get_ray_half_sep() const { return ray_half_sep_; }
	Complex<T> get_half_length_x() const { return half_length_x_; }
	Complex<T> get_center_x() const { return center_x_; }
	Complex<int> get_num_ray_threads() const { return num_ray_threads_; }
	int get_taylor_smooth() const { return taylor_smooth_; }
	T get_kappa_star_actual() const { return kappa_star_actual_; }
	T get_mean_mass() const { return mean_mass_; }
	T get_mean_mass2() const { return mean_mass2_; }
	T get_mean_mass_actual() const { return mean_mass_actual_; }
	T get_mean_mass2_actual() const { return mean_mass2_actual_; }
	T get_mean_mass2_ln_mass() const { return mean_mass2_ln_mass_; }
	T get_mean_mass2_ln_mass_actual() const { return mean_mass2_ln_mass_actual_; }
	T get_m_lower_actual() const { return m_lower_actual_; }
	T get_m_upper_actual() const { return m_upper_actual_; }
	int get_last_tile_rows() const { return engine_ ? engine_->last_tile_rows() : 0; }
	Engine* engine() { return engine_.get(); }
	bool is_built() const { return built_; }
	double mean_near_stars() const { return engine_ ? engine_->mean_near_stars(grid()) : 0; }
	const std::vector<long long>& histogram(int which, bool log_hist) const
	{
		return log_hist ? log_hist_[which] : hist_[which];
	}
	long long histogram_min(bool log_hist) const { return log_hist ? min_log_mag_ : min_mag_; }

	LensParams lens() const
	{
		LensParams L;
		L.kappa = double(kappa_tot);
		L.gamma = double(shear);
		L.theta = double(theta_star);
		L.kappastar = double(kappa_star);
		L.rectangular = rectangular;
		L.approx = approx;
		L.taylor_smooth = taylor_smooth_;
		L.corner_re = double(corner_.re);
		L.corner_im = double(corner_.im);
		return L;
	}

	GridParams grid() const
	{
		GridParams G;
		G.center_x[0] = double(center_x_.re);
		G.center_x[1] = double(center_x_.im);
		G.half_length_x[0] = double(half_length_x_.re);
		G.half_length_x[1] = double(half_length_x_.im);
		G.ray_half_sep[0] = double(ray_half_sep_.re);
		G.ray_half_sep[1] = double(ray_half_sep_.im);
		G.num_cells[0] = num_ray_threads_.re;
		G.num_cells[1] = num_ray_threads_.im;
		G.center_y[0] = double(center_y.re);
		G.center_y[1] = double(center_y.im);
		G.half_length_y[0] = double(half_length_y.re);
		G.half_length_y[1] = double(half_length_y.im);
		G.num_pixels[0] = num_pixels_y.re;
		G.num_pixels[1] = num_pixels_y.im;
		return G;
	}

private:
	const std::string outfile_type = ".bin";

	std::unique_ptr<Engine> engine_;
	bool prepared_ = false;
	bool built_ = false;

	int mf_id_ = MF_EQUAL;
	T mean_mass_ = 0, mean_mass2_ = 0, mean_mass2_ln_mass_ = 0;
	int num_stars_ = 0;
	T kappa_star_actual_ = 0, m_lower_actual_ = 0, m_upper_actual_ = 0;
	T mean_mass_actual_ = 0, mean_mass2_actual_ = 0, mean_mass2_ln_mass_actual_ = 0;
	T mu_ave_ = 0, num_rays_x_ = 0;
	Complex<T> ray_half_sep_, center_x_, half_length_x_, corner_;
	Complex<int> num_ray_threads_;
	int taylor_smooth_ = 1;
	T alpha_error_ = 0;
	int expansion_order_ = 0;
	T root_half_length_ = 0;
	int tree_levels_ = 0;
	double t_shoot_ = 0;
	long long cells_shot_ = 0;
	long long num_rays_received_ = 0;
	StarFileData file_stars_;

	std::vector<T> stars_host_;
	bool stars_stale_ = false;
	std::vector<T> pix_, pix_min_, pix_sad_;
	std::vector<int> ipix_, ipix_min_, ipix_sad_;
	long long min_mag_ = 0, max_mag_ = 0, min_log_mag_ = 0, max_log_mag_ = 0;
	std::vector<long long> hist_[3], log_hist_[3];

	/* ipm.cuh:235-368 */
	bool check_input_params(int verbose)
	{
		MLENS_VERBOSE("Checking input parameters...\n", verbose, 3);
		const T tmin = std::numeric_limits<T>::min();
		if (kappa_tot < tmin)
		{
			std::cerr << "Error. kappa_tot must be >= " << tmin << "\n";
			return false;
		}
		if (kappa_star < tmin)
		{
			std::cerr << "Error. kappa_star must be >= " << tmin << "\n";
			return false;
		}
		if (starfile == "" && kappa_star > kappa_tot)
		{
			std::cerr << "Error. kappa_star must be <= kappa_tot\n";
			return false;
		}
		if (starfile == "" && theta_star < tmin)
		{
			std::cerr << "Error. theta_star must be >= " << tmin << "\n";
			return false;
		}
		if (starfile == "" && mass_function_id(mass_function_str) == MF_INVALID)
		{
			std::cerr << "Error. mass_function must be equal, uniform, Salpeter, Kroupa, or optical_depth.\n";
			return false;
		}
		if (starfile == "" && m_solar < tmin)
		{
			std::cerr << "Error. m_solar must be >= " << tmin << "\n";
			return false;
		}
		if (starfile == "" && m_lower < tmin)
		{
			std::cerr << "Error. m_lower must be >= " << tmin << "\n";
			return false;
		}
		if (starfile == "" && m_upper < m_lower)
		{
			std::cerr << "Error. m_upper must be >= m_lower.\n";
			return false;
		}
		if (light_loss < tmin)
		{
			std::cerr << "Error. light_loss must be >= " << tmin << "\n";
			return false;
		}
		else if (light_loss > 0.01)
		{
			std::cerr << "Error. light_loss must be <= 0.01\n";
			return false;
		}
		if (starfile == "" && rectangular != 0 && rectangular != 1)
		{
			std::cerr << "Error. rectangular must be 1 (rectangular) or 0 (circular).\n";
			return false;
		}
		if (approx != 0 && approx != 1)
		{
			std::cerr << "Error. approx must be 1 (approximate) or 0 (exact).\n";
			return false;
		}
		if (safety_scale < 1.1)
		{
			std::cerr << "Error. safety_scale must be >= 1.1\n";
			return false;
		}
		if (half_length_y.re < tmin || half_length_y.im < tmin)
		{
			std::cerr << "Error. half_length_y1 and half_length_y2 must both be >= " << tmin << "\n";
			return false;
		}
		if (num_pixels_y.re < 1 || num_pixels_y.im < 1)
		{
			std::cerr << "Error. num_pixels_y1 and num_pixels_y2 must both be integers > 0\n";
			return false;
		}
		if (num_rays_y < 1)
		{
			std::cerr << "Error. num_rays_y must be an integer > 0\n";
			return false;
		}
		if (write_stars != 0 && write_stars != 1)
		{
			std::cerr << "Error. write_stars must be 1 (true) or 0 (false).\n";
			return false;
		}
		if (write_maps != 0 && write_maps != 1)
		{
			std::cerr << "Error. write_maps must be 1 (true) or 0 (false).\n";
			return false;
		}
		if (write_parities != 0 && write_parities != 1)
		{
			std::cerr << "Error. write_parities must be 1 (true) or 0 (false).\n";
			return false;
		}
		if (write_histograms != 0 && write_histograms != 1)
		{
			std::cerr << "Error. write_histograms must be 1 (true) or 0 (false).\n";
			return false;
		}
		if (expansion_order_override < 0 || expansion_order_override > MAX_EXPANSION_ORDER)
		{
			std::cerr << "Error. expansion_order must be 0 (automatic) or between 1 and " << MAX_EXPANSION_ORDER << "\n";
			return false;
		}
		if (m2l_mode < M2L_DIRECT || m2l_mode > M2L_MFMA_F64)
		{
			std::cerr << "Error. m2l_mode must be 0 (direct FP64), 1 (MFMA, FP32 for p <= 20), 2 (FP32 MFMA) or 3 (FP64 MFMA).\n";
			return false;
		}
		if (tile_rows < 0 || tile_rows > 65536)
		{
			std::cerr << "Error. tile_rows must be 0 (chosen per launch) or between 1 and 65536.\n";
			return false;
		}
		if (kernel != KERNEL_FAST && kernel != KERNEL_REFERENCE)
		{
			std::cerr << "Error. kernel must be 0 (fast) or 1 (reference).\n";
			return false;
		}
		if (precision != PREC_FP64 && precision != PREC_MIXED)
		{
			std::cerr << "Error. precision must be 0 (fp64) or 1 (mixed).\n";
			return false;
		}
		if (kernel == KERNEL_REFERENCE && (!IS_IPM || precision != PREC_FP64))
		{
			std::cerr << "Error. the reference kernel exists for IPM in fp64 only.\n";
			return false;
		}
		MLENS_VERBOSE("Done checking input parameters.\n\n", verbose, 3);
		return true;
	}

	void star_params_from(const std::vector<double>& aos, int n)
	{
		// star.cuh:124-161
		double mt = 0, m2 = 0, m2l = 0;
		T lo = std::numeric_limits<T>::max(), hi = std::numeric_limits<T>::min();
		for (int i = 0; i < n; i++)
		{
			T m = T(aos[3 * i + 2]);
			mt += m;
			m2 += m * m;
			m2l += m * m * std::log(m);
			lo = std::min(lo, m);
			hi = std::max(hi, m);
		}
		apply_star_totals(mt, m2, m2l, lo, hi, n);
	}

	void apply_star_totals(double mt, double m2, double m2l, T lo, T hi, int n)
	{
		m_lower_actual_ = lo;
		m_upper_actual_ = hi;
		mean_mass_actual_ = T(mt / n);
		mean_mass2_actual_ = T(m2 / n);
		mean_mass2_ln_mass_actual_ = T(m2l / n);
		if (rectangular)
			kappa_star_actual_ = T(mt * std::numbers::pi_v<T> * theta_star * theta_star / (4 * corner_.re * corner_.im));
		else
			kappa_star_actual_ = T(mt * theta_star * theta_star / (corner_.abs() * corner_.abs()));
	}

	/* ipm.cuh:370-565 */
	bool calculate_derived_params(int verbose)
	{
		MLENS_VERBOSE("Calculating derived parameters...\n", verbose, 3);
		if (starfile == "")
		{
			if (mass_function_str == "equal")
			{
				m_lower = 1;
				m_upper = 1;
			}
			else
			{
				m_lower = m_lower * m_solar;
				m_upper = m_upper * m_solar;
			}
			mf_id_ = mass_function_id(mass_function_str);
			T mm[3];
			mass_moments<T>(mf_id_, m_lower, m_upper, m_solar, mm);
			mean_mass_ = mm[0];
			mean_mass2_ = mm[1];
			mean_mass2_ln_mass_ = mm[2];
		}
		else
		{
			if (!read_star_file(starfile, file_stars_))
			{
				std::cerr << "Error. Unable to read star field parameters from file " << starfile << "\n";
				return false;
			}
			num_stars_ = file_stars_.num_stars;
			rectangular = file_stars_.rectangular;
			corner_ = Complex<T>(T(file_stars_.corner_re), T(file_stars_.corner_im));
			theta_star = T(file_stars_.theta);
			star_params_from(file_stars_.stars, num_stars_);
			kappa_star = kappa_star_actual_;
			m_lower = m_lower_actual_;
			m_upper = m_upper_actual_;
			mean_mass_ = mean_mass_actual_;
			mean_mass2_ = mean_mass2_actual_;
			mean_mass2_ln_mass_ = mean_mass2_ln_mass_actual_;
			if (kappa_star > kappa_tot) std::cerr << "Warning. kappa_star > kappa_tot\n";
		}

		mu_ave_ = T(1 / ((1 - kappa_tot) * (1 - kappa_tot) - shear * shear));
		num_rays_x_ = T(1.0 * num_rays_y * num_pixels_y.re * num_pixels_y.im / (2 * half_length_y.re * 2 * half_length_y.im));
		ray_half_sep_ = Complex<T>(std::sqrt(half_length_y.re / half_length_y.im * num_pixels_y.im / num_pixels_y.re),
			std::sqrt(half_length_y.im / half_length_y.re * num_pixels_y.re / num_pixels_y.im));
		{
			T d = 2 * std::sqrt(num_rays_x_);
			ray_half_sep_ = Complex<T>(ray_half_sep_.re / d, ray_half_sep_.im / d);
		}
		T ext = theta_star * std::sqrt(kappa_star * mean_mass2_ / (mean_mass_ * light_loss));
		half_length_x_ = Complex<T>(half_length_y.re + ext, half_length_y.im + ext);
		half_length_x_ = Complex<T>(half_length_x_.re / std::abs(1 - kappa_tot + shear), half_length_x_.im / std::abs(1 - kappa_tot - shear));
		num_ray_threads_ = Complex<int>(static_cast<int>(half_length_x_.re / (2 * ray_half_sep_.re)) + 1,
			static_cast<int>(half_length_x_.im / (2 * ray_half_sep_.im)) + 1);
		half_length_x_ = Complex<T>(2 * ray_half_sep_.re * num_ray_threads_.re, 2 * ray_half_sep_.im * num_ray_threads_.im);
		num_ray_threads_ = Complex<int>(2 * num_ray_threads_.re, 2 * num_ray_threads_.im);
		center_x_ = Complex<T>(center_y.re / (1 - kappa_tot + shear), center_y.im / (1 - kappa_tot - shear));

		if (starfile == "")
		{
			corner_ = Complex<T>(std::abs(center_x_.re) + half_length_x_.re, std::abs(center_x_.im) + half_length_x_.im);
			if (rectangular)
			{
				num_stars_ = static_cast<int>(std::ceil(safety_scale * 2 * corner_.re * safety_scale * 2 * corner_.im * kappa_star /
					(std::numbers::pi_v<T> * theta_star * theta_star * mean_mass_)));
				corner_ = Complex<T>(std::sqrt(corner_.re / corner_.im), std::sqrt(corner_.im / corner_.re));
				T f = std::sqrt(std::numbers::pi_v<T> * theta_star * theta_star * num_stars_ * mean_mass_ / (4 * kappa_star));
				corner_ = Complex<T>(corner_.re * f, corner_.im * f);
			}
			else
			{
				num_stars_ = static_cast<int>(std::ceil(safety_scale * corner_.abs() * safety_scale * corner_.abs() * kappa_star /
					(theta_star * theta_star * mean_mass_)));
				T a = corner_.abs();
				corner_ = Complex<T>(corner_.re / a, corner_.im / a);
				T f = std::sqrt(theta_star * theta_star * num_stars_ * mean_mass_ / kappa_star);
				corner_ = Complex<T>(corner_.re * f, corner_.im * f);
			}
		}
		else
		{
			Complex<T> tmp = Complex<T>(std::abs(center_x_.re) + half_length_x_.re, std::abs(center_x_.im) + half_length_x_.im);
			if (!rectangular)
			{
				T a = tmp.abs(), c = corner_.abs();
				corner_ = Complex<T>(tmp.re / a * c, tmp.im / a * c);
			}
			if ((rectangular && (corner_.re < safety_scale * tmp.re || corner_.im < safety_scale * tmp.im)) ||
				(!rectangular && (corner_.abs() < safety_scale * tmp.abs())))
			{
				std::cerr << "Error. The provided star field is not large enough to cover the desired source plane region.\n";
				std::cerr << "Try decreasing the safety_scale, or providing a larger field of stars.\n";
				return false;
			}
		}
		if (num_stars_ < 1)
		{
			std::cerr << "Error. The derived number of stars must be >= 1\n";
			return false;
		}

		alpha_error_ = std::min(half_length_y.re / (10 * num_pixels_y.re), half_length_y.im / (10 * num_pixels_y.im));
		taylor_smooth_ = 1;
		while ((kappa_star * std::numbers::inv_pi_v<T> * 4 / (taylor_smooth_ + 1) * corner_.abs() * (safety_scale + 1) /
					   (safety_scale - 1) * std::pow(1 / safety_scale, taylor_smooth_ + 1) >
				   alpha_error_) &&
			taylor_smooth_ <= MAX_TAYLOR_SMOOTH)
		{
			taylor_smooth_ += 2;
		}
		while ((std::fmod(corner_.arg() * (taylor_smooth_ - 1), std::numbers::pi_v<T>) < 0.1 * std::numbers::pi_v<T> ||
				   std::fmod(corner_.arg() * (taylor_smooth_ - 1), std::numbers::pi_v<T>) > 0.9 * std::numbers::pi_v<T>) &&
			taylor_smooth_ <= MAX_TAYLOR_SMOOTH)
		{
			taylor_smooth_ += 2;
		}
		if (rectangular && taylor_smooth_ > MAX_TAYLOR_SMOOTH)
		{
			std::cerr << "Error. taylor_smooth must be <= " << MAX_TAYLOR_SMOOTH << "\n";
			return false;
		}
		MLENS_VERBOSE("Done calculating derived parameters.\n\n", verbose, 3);
		return true;
	}

	bool set_device(int verbose)
	{
		int n = 0;
		hipError_t e = hipGetDeviceCount(&n);
		if (e != hipSuccess || n < 1)
		{
			std::cerr << "Error. No HIP capable devices detected.\n";
			return false;
		}
		MLENS_HIP_CHECK(hipSetDevice(0));
		if (!engine_) engine_ = std::make_unique<Engine>();
		engine_->alloc_maps(num_pixels_y.re, num_pixels_y.im, IS_IPM, Precision(precision), write_parities != 0);
		return true;
	}

	/* ipm.cuh:633-738 */
	bool populate_star_array(int verbose)
	{
		if (starfile == "")
		{
			MLENS_VERBOSE("Generating star field...\n", verbose, 1);
			while (random_seed == 0)
				random_seed = static_cast<int>(std::chrono::system_clock::now().time_since_epoch().count());
			engine_->generate_stars(num_stars_, rectangular, double(corner_.re), double(corner_.im), mf_id_, double(m_lower),
				double(m_upper), double(m_solar), uint32_t(random_seed));
			StarStats s = engine_->star_stats();
			apply_star_totals(s.m_tot, s.m2_tot, s.m2_ln_tot, T(s.m_min), T(s.m_max), num_stars_);
			// corner re-scaled with the realised mean mass; stars are not moved (ipm.cuh:717-731)
			if (rectangular)
			{
				corner_ = Complex<T>(std::sqrt(corner_.re / corner_.im), std::sqrt(corner_.im / corner_.re));
				T f = std::sqrt(std::numbers::pi_v<T> * theta_star * theta_star * num_stars_ * mean_mass_actual_ / (4 * kappa_star));
				corner_ = Complex<T>(corner_.re * f, corner_.im * f);
			}
			else
			{
				T a = corner_.abs();
				corner_ = Complex<T>(corner_.re / a, corner_.im / a);
				T f = std::sqrt(theta_star * theta_star * num_stars_ * mean_mass_actual_ / kappa_star);
				corner_ = Complex<T>(corner_.re * f, corner_.im * f);
			}
			MLENS_VERBOSE("Done generating star field.\n\n", verbose, 1);
		}
		else
		{
			random_seed = 0;
			engine_->upload_stars(file_stars_.stars.data(), num_stars_);
			star_params_from(file_stars_.stars, num_stars_);
		}
		return true;
	}

	/* ipm.cuh:740-912 */
	bool create_tree(int verbose)
	{
		if (rectangular)
			root_half_length_ = corner_.re > corner_.im ? corner_.re : corner_.im;
		else
			root_half_length_ = corner_.abs();
		root_half_length_ = static_cast<T>(root_half_length_ * 1.1);
		MLENS_VERBOSE("Creating children and sorting stars...\n", verbose, 1);
		engine_->build_tree(double(root_half_length_));
		tree_levels_ = engine_->tree_levels();
		MLENS_VERBOSE("Done creating children and sorting stars.\n\n", verbose, 1);
		if (expansion_order_override > 0)
		{
			expansion_order_ = expansion_order_override;
		}
		else
		{
			expansion_order_ = static_cast<int>(std::ceil(2 * std::log2(theta_star) + tree_levels_ - std::log2(root_half_length_) -
				std::log2(alpha_error_)));
			if (expansion_order_ < 3)
			{
				std::cerr << "Error. Expansion order needs to be >= 3\n";
				return false;
			}
			else if (expansion_order_ > MAX_EXPANSION_ORDER)
			{
				std::cerr << "Error. Maximum allowed expansion order is " << MAX_EXPANSION_ORDER << "\n";
				return false;
			}
		}
		MLENS_VERBOSE("Calculating multipole and local coefficients...\n", verbose, 1);
		engine_->set_m2l_mode(m2l_mode);
		engine_->fmm(expansion_order_, lens());
		MLENS_VERBOSE("Done calculating multipole and local coefficients.\n\n", verbose, 1);
		return true;
	}

	/* ipm.cuh:950-1120 and irs.cuh:941- */
	bool create_histograms(int verbose)
	{
		for (int k = 0; k < 3; k++)
		{
			hist_[k].clear();
			log_hist_[k].clear();
		}
		if (!write_histograms) return true;
		const int nmaps = write_parities ? 3 : 1;
		const int factor = IS_IPM ? 1000 : 1;
		T mu_min_theory = 1 / ((1 - kappa_tot + kappa_star) * (1 - kappa_tot + kappa_star));
		long long lo = std::numeric_limits<long long>::max(), hi = std::numeric_limits<long long>::min();
		for (int k = 0; k < nmaps; k++)
		{
			double mn, mx;
			engine_->map_minmax(k, &mn, &mx);
			// in double, as k_histogram bins every pixel (the reference rounds these two in T
			// and its pixels in double, ipm.cuh:972-973 and array_functions.cuh:127, so a float
			// map's extreme pixel can fall outside its own histogram there)
			long long a = IS_IPM ? std::llround(double(T(mn)) * factor) : (long long)mn;
			long long b = IS_IPM ? std::llround(double(T(mx)) * factor) : (long long)mx;
			T mu_min_actual = IS_IPM ? T(1.0 * a / factor) : T(1.0 * a / num_rays_y);
			if (k < 2 && mu_ave_ > 1 && mu_min_actual < mu_min_theory)
			{
				std::cerr << (k == 0 ? "Warning. Minimum magnification after shooting is less than the theoretical minimum.\n"
									 : "Warning. Minimum positive parity magnification after shooting is less than the theoretical minimum.\n");
				std::cerr << "   mu_min_actual = " << mu_min_actual << "\n";
				std::cerr << "   mu_min_theory = 1 / (1 - (kappa_tot - kappa_star))^2 = " << mu_min_theory << "\n";
			}
			lo = std::min(lo, a);
			hi = std::max(hi, b);
		}
		min_mag_ = lo;
		max_mag_ = hi;
		long long len = hi - lo + 1;
		for (int k = 0; k < nmaps; k++)
		{
			hist_[k].assign(len, 0);
			engine_->map_histogram(k, false, factor, lo, len, hist_[k].data());
		}
		if (IS_IPM)
		{
			// log10 histogram over the positive pixels (a zero pixel has no log10)
			long long llo = std::numeric_limits<long long>::max(), lhi = std::numeric_limits<long long>::min();
			for (int k = 0; k < nmaps; k++)
			{
				double mn, mx;
				engine_->map_minmax(k, &mn, &mx);
				double pos_min = mn > 0 ? mn : positive_min(k);
				if (!(pos_min > 0) || !(mx > 0)) continue;
				llo = std::min(llo, (long long)std::llround(std::log10(double(T(pos_min))) * factor));
				lhi = std::max(lhi, (long long)std::llround(std::log10(double(T(mx))) * factor));
			}
			if (lhi >= llo)
			{
				min_log_mag_ = llo;
				max_log_mag_ = lhi;
				for (int k = 0; k < nmaps; k++)
				{
					log_hist_[k].assign(lhi - llo + 1, 0);
					engine_->map_histogram(k, true, factor, llo, lhi - llo + 1, log_hist_[k].data());
				}
			}
		}
		return true;
	}

	double positive_min(int k)
	{
		// smallest positive pixel of map k, on the host copy of the map
		size_t n = size_t(num_pixels_y.re) * num_pixels_y.im;
		std::vector<double> v(n);
		if (engine_->maps_are_fp32())
		{
			std::vector<float> f(n);
			engine_->download_map(k, f.data());
			for (size_t i = 0; i < n; i++) v[i] = f[i];
		}
		else
		{
			engine_->download_map(k, v.data());
		}
		double m = std::numeric_limits<double>::infinity();
		for (double x : v)
			if (x > 0 && x < m) m = x;
		return m;
	}

	// the star field in tree order, as get_stars returns it
	void download_stars()
	{
		std::vector<double> s(size_t(num_stars_) * 3);
		engine_->download_stars(s.data());
		stars_host_.assign(s.begin(), s.end());
		stars_stale_ = false;
	}

	void download_results()
	{
		size_t n = size_t(num_pixels_y.re) * num_pixels_y.im;
		auto fetch = [&](int which, std::vector<T>& dst) {
			dst.resize(n);
			if (engine_->maps_are_fp32())
			{
				if (std::is_same<T, float>::value)
					engine_->download_map(which, dst.data());
				else
				{
					std::vector<float> f(n);
					engine_->download_map(which, f.data());
					for (size_t i = 0; i < n; i++) dst[i] = T(f[i]);
				}
			}
			else
			{
				if (std::is_same<T, double>::value)
					engine_->download_map(which, dst.data());
				else
				{
					std::vector<double> d(n);
					engine_->download_map(which, d.data());
					for (size_t i = 0; i < n; i++) dst[i] = T(d[i]);
				}
			}
		};
		if (IS_IPM)
		{
			fetch(0, pix_);
			if (write_parities)
			{
				fetch(1, pix_min_);
				fetch(2, pix_sad_);
			}
			else
			{
				pix_min_.clear();
				pix_sad_.clear();
			}
		}
		else
		{
			ipix_.resize(n);
			engine_->download_map(0, ipix_.data());
			if (write_parities)
			{
				ipix_min_.resize(n);
				ipix_sad_.resize(n);
				engine_->download_map(1, ipix_min_.data());
				engine_->download_map(2, ipix_sad_.data());
			}
			else
			{
				ipix_min_.clear();
				ipix_sad_.clear();
			}
		}
		download_stars();
	}

	/* ipm.cuh:1122-1322 */
	bool write_files(int verbose)
	{
		const std::string tag = IS_IPM ? "ipm" : "irs";
		std::string fname = outfile_prefix + tag + "_parameter_info.txt";
		std::ofstream out(fname);
		if (!out.is_open())
		{
			std::cerr << "Error. Failed to open file " << fname << "\n";
			return false;
		}
		out.precision(9);
		out << "kappa_tot " << kappa_tot << "\n";
		out << "shear " << shear << "\n";
		out << "mu_ave " << mu_ave_ << "\n";
		out << "smooth_fraction " << (1 - kappa_star / kappa_tot) << "\n";
		out << "kappa_star " << kappa_star << "\n";
		if (starfile == "") out << "kappa_star_actual " << kappa_star_actual_ << "\n";
		out << "theta_star " << theta_star << "\n";
		out << "random_seed " << random_seed << "\n";
		if (starfile == "")
		{
			out << "mass_function " << mass_function_str << "\n";
			if (mass_function_str == "salpeter" || mass_function_str == "kroupa") out << "m_solar " << m_solar << "\n";
			out << "m_lower " << m_lower << "\n";
			out << "m_upper " << m_upper << "\n";
			out << "mean_mass " << mean_mass_ << "\n";
			out << "mean_mass2 " << mean_mass2_ << "\n";
			out << "mean_mass2_ln_mass " << mean_mass2_ln_mass_ << "\n";
		}
		out << "m_lower_actual " << m_lower_actual_ << "\n";
		out << "m_upper_actual " << m_upper_actual_ << "\n";
		out << "mean_mass_actual " << mean_mass_actual_ << "\n";
		out << "mean_mass2_actual " << mean_mass2_actual_ << "\n";
		out << "mean_mass2_ln_mass_actual " << mean_mass2_ln_mass_actual_ << "\n";
		out << "light_loss " << light_loss << "\n";
		out << "num_stars " << num_stars_ << "\n";
		if (rectangular)
		{
			out << "corner_x1 " << corner_.re << "\n";
			out << "corner_x2 " << corner_.im << "\n";
			if (approx) out << "taylor_smooth " << taylor_smooth_ << "\n";
		}
		else
		{
			out << "rad " << corner_.abs() << "\n";
		}
		out << "safety_scale " << safety_scale << "\n";
		out << "center_y1 " << center_y.re << "\n";
		out << "center_y2 " << center_y.im << "\n";
		out << "half_length_y1 " << half_length_y.re << "\n";
		out << "half_length_y2 " << half_length_y.im << "\n";
		out << "num_pixels_y1 " << num_pixels_y.re << "\n";
		out << "num_pixels_y2 " << num_pixels_y.im << "\n";
		out << "center_x1 " << center_x_.re << "\n";
		out << "center_x2 " << center_x_.im << "\n";
		out << "half_length_x1 " << half_length_x_.re << "\n";
		out << "half_length_x2 " << half_length_x_.im << "\n";
		out << "num_rays_y " << num_rays_y << "\n";
		if (!IS_IPM) out << "mean_num_rays_y " << (1.0 * num_rays_received_ / (double(num_pixels_y.re) * num_pixels_y.im)) << "\n";
		out << "num_rays_x " << num_rays_x_ << "\n";
		out << "ray_half_sep_1 " << ray_half_sep_.re << "\n";
		out << "ray_half_sep_2 " << ray_half_sep_.im << "\n";
		out << "alpha_error " << alpha_error_ << "\n";
		out << "expansion_order " << expansion_order_ << "\n";
		out << "root_half_length " << root_half_length_ << "\n";
		out << "num_ray_threads_1 " << num_ray_threads_.re << "\n";
		out << "num_ray_threads_2 " << num_ray_threads_.im << "\n";
		out << "tree_levels " << tree_levels_ << "\n";
		if (IS_IPM)
			out << "t_shoot_cells " << t_shoot_ << "\n";
		else
		{
			out << "t_shoot_rays " << t_shoot_ << "\n";
			out << "num_rays_shot " << get_num_rays_shot() << "\n";
			out << "num_rays_received " << num_rays_received_ << "\n";
		}
		out.close();
		MLENS_VERBOSE("Done writing parameter info to file " << fname << "\n", verbose, 1);

		if (write_stars)
		{
			fname = outfile_prefix + tag + "_stars" + outfile_type;
			if (!write_star_file<T>(fname, num_stars_, rectangular, corner_, theta_star, stars_host_.data()))
			{
				std::cerr << "Error. Unable to write star info to file " << fname << "\n";
				return false;
			}
		}
		if (write_histograms)
		{
			const char* suffix[3] = {"", "_minima", "_saddles"};
			const int nmaps = write_parities ? 3 : 1;
			for (int k = 0; k < nmaps; k++)
			{
				fname = outfile_prefix + tag + (IS_IPM ? "_mags_numpixels" : "_numrays_numpixels") + suffix[k] + ".txt";
				if (!write_histogram(fname, hist_[k], min_mag_))
				{
					std::cerr << "Error. Unable to write magnification histogram to file " << fname << "\n";
					return false;
				}
				if (IS_IPM)
				{
					fname = outfile_prefix + tag + "_log_mags_numpixels" + suffix[k] + ".txt";
					if (!write_histogram(fname, log_hist_[k], min_log_mag_))
					{
						std::cerr << "Error. Unable to write magnification histogram to file " << fname << "\n";
						return false;
					}
				}
			}
		}
		if (write_maps)
		{
			const char* suffix[3] = {"", "_minima", "_saddles"};
			const int nmaps = write_parities ? 3 : 1;
			for (int k = 0; k < nmaps; k++)
			{
				fname = outfile_prefix + tag + "_magnifications" + suffix[k] + outfile_type;
				bool ok;
				if (IS_IPM)
					ok = write_array<T>(fname, (k == 0 ? pix_ : (k == 1 ? pix_min_ : pix_sad_)).data(), num_pixels_y.im, num_pixels_y.re);
				else
					ok = write_array<int>(fname, (k == 0 ? ipix_ : (k == 1 ? ipix_min_ : ipix_sad_)).data(), num_pixels_y.im, num_pixels_y.re);
				if (!ok)
				{
					std::cerr << "Error. Unable to write magnifications to file " << fname << "\n";
					return false;
				}
			}
		}
		return true;
	}
};

}  // namespace mlens
