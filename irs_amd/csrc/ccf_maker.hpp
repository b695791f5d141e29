// Host orchestration of the critical curves and caustics of a star field: the
// CCF<T> class of the reference (ccf.cuh:34-1298) with the same public fields
// and defaults, the same validation and derived parameters, and run() / save().
// As for the maps: validation and derivation come before any device call, the
// device is not reset, returned arrays are host copies valid until this
// object's next run(), and the phases are timed with HIP events (seconds).
// The roots are refined in FP64 (the library instantiates T = double).
#pragma once

#include "common.hpp"
#include "engine.hpp"
#include "io.hpp"
#include "mass_functions.hpp"

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

template <typename T>
class CCFMaker
{
public:
	/* public input fields and defaults (ccf.cuh:42-62) */
	T kappa_tot = static_cast<T>(0.3);
	T shear = static_cast<T>(0.3);
	T kappa_star = static_cast<T>(0.27);
	T theta_star = static_cast<T>(1);
	std::string mass_function_str = "equal";
	T m_solar = static_cast<T>(1);
	T m_lower = static_cast<T>(0.01);
	T m_upper = static_cast<T>(50);
	int rectangular = 0;
	int approx = 1;
	T safety_scale = static_cast<T>(1.37);
	int num_stars = 137;
	std::string starfile = "";
	int num_phi = 100;
	int num_branches = 1;
	int random_seed = 0;
	int write_stars = 1;
	int write_critical_curves = 1;
	int write_caustics = 1;
	int write_mu_length_scales = 0;
	std::string outfile_prefix = "./";

	/* ccf.cuh:1263-1280 */
	bool run(int verbose)
	{
		try
		{
			if (!check_input_params(verbose)) return false;
			if (!calculate_derived_params(verbose)) return false;
			if (!set_device()) return false;
			if (!populate_star_array(verbose)) return false;
			if (!create_tree(verbose)) return false;
			return find_ccs(verbose);
		}
		catch (const HipError& e)
		{
			std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";
			return false;
		}
	}

	bool save(int verbose) { return write_files(verbose); }

	int get_num_roots() const { return num_roots_; }
	Complex<T> get_corner() const { return rectangular ? corner_ : Complex<T>(corner_.abs(), 0); }
	T* get_stars() { return stars_host_.empty() ? nullptr : stars_host_.data(); }
	T* get_critical_curves() { return ccs_.empty() ? nullptr : ccs_.data(); }
	T* get_caustics() { return caustics_.empty() ? nullptr : caustics_.data(); }
	T* get_mu_length_scales() { return mu_.empty() ? nullptr : mu_.data(); }
	double get_t_ccs() const { return t_ccs_; }
	double get_t_init_roots() const { return t_init_roots_; }
	double get_max_error() const { return max_error_; }
	int get_expansion_order() const { return expansion_order_; }
	int get_tree_levels() const { return tree_levels_; }
	int get_taylor_smooth() const { return taylor_smooth_; }

private:
	const std::string outfile_type = ".bin";
	static constexpr int MAX_TAYLOR_SMOOTH_CCF = 101;

	std::unique_ptr<Engine> engine_;
	int mf_id_ = MF_EQUAL;
	T mean_mass_ = 0, mean_mass2_ = 0, mean_mass2_ln_mass_ = 0;
	T kappa_star_actual_ = 0, m_lower_actual_ = 0, m_upper_actual_ = 0;
	T mean_mass_actual_ = 0, mean_mass2_actual_ = 0, mean_mass2_ln_mass_actual_ = 0;
	T mu_ave_ = 0;
	Complex<T> corner_;
	int taylor_smooth_ = 1;
	T alpha_error_ = 0;
	int expansion_order_ = 0;
	T root_half_length_ = 0;
	int tree_levels_ = 0;
	int num_roots_ = 0;
	double max_error_ = 0, t_init_roots_ = 0, t_ccs_ = 0, t_caustics_ = 0;
	StarFileData file_stars_;
	std::vector<double> init_stars_;   // AoS x1, x2, m in generation (or file) order
	std::vector<T> stars_host_, ccs_, caustics_, mu_;

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

	/* ccf.cuh:226-348 */
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
		if (num_stars < 1)
		{
			std::cerr << "Error. num_stars must be an integer > 0\n";
			return false;
		}
		if (num_phi < 1 || num_phi % 2 != 0)
		{
			std::cerr << "Error. num_phi must be an even integer > 0\n";
			return false;
		}
		if (num_branches < 1)
		{
			std::cerr << "Error. num_branches must be an integer > 0\n";
			return false;
		}
		if (num_phi % (2 * num_branches) != 0)
		{
			std::cerr << "Error. num_phi must be a multiple of 2 * num_branches\n";
			return false;
		}
		const int* flags[4] = {&write_stars, &write_critical_curves, &write_caustics, &write_mu_length_scales};
		const char* names[4] = {"write_stars", "write_critical_curves", "write_caustics", "write_mu_length_scales"};
		for (int k = 0; k < 4; k++)
		{
			if (*flags[k] != 0 && *flags[k] != 1)
			{
				std::cerr << "Error. " << names[k] << " must be 1 (true) or 0 (false).\n";
				return false;
			}
		}
		MLENS_VERBOSE("Done checking input parameters.\n\n", verbose, 3);
		return true;
	}

	/* ccf.cuh:350-486 */
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
			num_stars = file_stars_.num_stars;
			rectangular = file_stars_.rectangular;
			corner_ = Complex<T>(T(file_stars_.corner_re), T(file_stars_.corner_im));
			theta_star = T(file_stars_.theta);
			star_params_from(file_stars_.stars, num_stars);
			kappa_star = kappa_star_actual_;
			if (kappa_star > kappa_tot) std::cerr << "Warning. kappa_star > kappa_tot\n";
			m_lower = m_lower_actual_;
			m_upper = m_upper_actual_;
			mean_mass_ = mean_mass_actual_;
			mean_mass2_ = mean_mass2_actual_;
			mean_mass2_ln_mass_ = mean_mass2_ln_mass_actual_;
		}
		mu_ave_ = T(1 / ((1 - kappa_tot) * (1 - kappa_tot) - shear * shear));
		if (starfile == "") corner_ = field_corner(mean_mass_);
		// 10^-7 Einstein radii
		alpha_error_ = static_cast<T>(theta_star * 0.0000001);
		taylor_smooth_ = 1;
		while ((kappa_star * std::numbers::inv_pi_v<T> * 4 / (taylor_smooth_ + 1) * corner_.abs() * (safety_scale + 1) /
					   (safety_scale - 1) * std::pow(1 / safety_scale, taylor_smooth_ + 1) >
				   alpha_error_) &&
			taylor_smooth_ <= MAX_TAYLOR_SMOOTH_CCF)
		{
			taylor_smooth_ += 2;
		}
		while ((std::fmod(corner_.arg() * (taylor_smooth_ - 1), std::numbers::pi_v<T>) < 0.1 * std::numbers::pi_v<T> ||
				   std::fmod(corner_.arg() * (taylor_smooth_ - 1), std::numbers::pi_v<T>) > 0.9 * std::numbers::pi_v<T>) &&
			taylor_smooth_ <= MAX_TAYLOR_SMOOTH_CCF)
		{
			taylor_smooth_ += 2;
		}
		if (rectangular && taylor_smooth_ > MAX_TAYLOR_SMOOTH_CCF)
		{
			std::cerr << "Error. taylor_smooth must be <= " << MAX_TAYLOR_SMOOTH_CCF << "\n";
			return false;
		}
		// ccf.cuh:463-483: two roots per star, plus the smooth matter's own
		if (rectangular)
			num_roots_ = approx ? 2 * num_stars + taylor_smooth_ - 1 : 2 * num_stars;
		else
			num_roots_ = approx ? 2 * num_stars : 2 * num_stars + 2;
		MLENS_VERBOSE("Done calculating derived parameters.\n\n", verbose, 3);
		return true;
	}

	// ccf.cuh:429-446 and 671-685: field size from the number of stars and their mean mass
	Complex<T> field_corner(T mean_mass) const
	{
		Complex<T> c(1 / std::abs(1 - kappa_tot + shear), 1 / std::abs(1 - kappa_tot - shear));
		if (rectangular)
		{
			c = Complex<T>(std::sqrt(c.re / c.im), std::sqrt(c.im / c.re));
			T f = std::sqrt(std::numbers::pi_v<T> * theta_star * theta_star * num_stars * mean_mass / (4 * kappa_star));
			return Complex<T>(c.re * f, c.im * f);
		}
		T a = c.abs();
		T f = std::sqrt(theta_star * theta_star * num_stars * mean_mass / kappa_star);
		return Complex<T>(c.re / a * f, c.im / a * f);
	}

	bool set_device()
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
		return true;
	}

	// star.cuh:124-161
	void star_params_from(const std::vector<double>& aos, int n)
	{
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

	/* ccf.cuh:596-705 */
	bool populate_star_array(int verbose)
	{
		init_stars_.assign(3 * size_t(num_stars), 0.0);
		if (starfile == "")
		{
			MLENS_VERBOSE("Generating star field...\n", verbose, 1);
			while (random_seed == 0)
				random_seed = static_cast<int>(std::chrono::system_clock::now().time_since_epoch().count());
			engine_->generate_stars(num_stars, rectangular, double(corner_.re), double(corner_.im), mf_id_, double(m_lower),
				double(m_upper), double(m_solar), uint32_t(random_seed));
			engine_->download_stars(init_stars_.data());
			star_params_from(init_stars_, num_stars);
			// corner re-scaled with the realised mean mass; stars are not moved
			corner_ = field_corner(mean_mass_actual_);
			MLENS_VERBOSE("Done generating star field.\n\n", verbose, 1);
		}
		else
		{
			random_seed = 0;
			init_stars_ = file_stars_.stars;
			engine_->upload_stars(init_stars_.data(), num_stars);
		}
		return true;
	}

	/* ccf.cuh:707-904 */
	bool create_tree(int verbose)
	{
		root_half_length_ = rectangular ? (corner_.re > corner_.im ? corner_.re : corner_.im) : corner_.abs();
		root_half_length_ = static_cast<T>(root_half_length_ * 1.1);
		MLENS_VERBOSE("Creating children and sorting stars...\n", verbose, 1);
		engine_->build_tree(double(root_half_length_));
		tree_levels_ = engine_->tree_levels();
		expansion_order_ = static_cast<int>(
			std::ceil(2 * std::log2(theta_star) + tree_levels_ - std::log2(root_half_length_) - std::log2(alpha_error_)));
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
		MLENS_VERBOSE("Calculating multipole and local coefficients...\n", verbose, 1);
		engine_->set_m2l_mode(M2L_MFMA_F64);
		engine_->fmm(expansion_order_, lens());
		return true;
	}

	/* ccf.cuh:906-1100 */
	bool find_ccs(int verbose)
	{
		// starting roots at every branch centre: each star +-1 on the real axis, and
		// the smooth matter's roots evenly on the circle of the field's radius
		std::vector<double> init(2 * size_t(num_branches) * num_roots_);
		const int extra = num_roots_ - 2 * num_stars;
		for (int c = 0; c < num_branches; c++)
		{
			double* r = init.data() + 2 * size_t(c) * num_roots_;
			for (int i = 0; i < num_stars; i++)
			{
				r[2 * i] = init_stars_[3 * i] + 1;
				r[2 * i + 1] = init_stars_[3 * i + 1];
				r[2 * (num_stars + i)] = init_stars_[3 * i] - 1;
				r[2 * (num_stars + i) + 1] = init_stars_[3 * i + 1];
			}
			for (int i = 0; i < extra; i++)
			{
				const double t = 2 * std::numbers::pi / extra * i;
				r[2 * (2 * num_stars + i)] = double(corner_.abs()) * std::cos(t);
				r[2 * (2 * num_stars + i) + 1] = double(corner_.abs()) * std::sin(t);
			}
		}
		const size_t total = size_t(num_phi + num_branches) * num_roots_;
		std::vector<double> ccs(2 * total), caust(write_caustics ? 2 * total : 0), mu(write_mu_length_scales ? total : 0);
		MLENS_VERBOSE("Finding critical curve positions...\n", verbose, 1);
		CCFResult r = engine_->ccf_solve(lens(), init.data(), num_roots_, num_phi, num_branches, !rectangular && !approx, ccs.data(),
			write_caustics ? caust.data() : nullptr, write_mu_length_scales ? mu.data() : nullptr);
		t_init_roots_ = r.t_init_ms / 1000.0;
		t_ccs_ = r.t_ccs_ms / 1000.0;
		t_caustics_ = r.t_caustics_ms / 1000.0;
		max_error_ = r.max_error;
		if (r.has_nan)
		{
			std::cerr << "Error. Errors in 1/mu contain values which are not positive real numbers.\n";
			return false;
		}
		MLENS_VERBOSE("Done finding critical curve positions. Elapsed time: " << t_ccs_ << " seconds.\n", verbose, 1);
		MLENS_VERBOSE("Maximum error in 1/mu: " << max_error_ << "\n\n", verbose, 1);
		ccs_.assign(ccs.begin(), ccs.end());
		caustics_.assign(caust.begin(), caust.end());
		mu_.assign(mu.begin(), mu.end());
		std::vector<double> st(3 * size_t(num_stars));
		engine_->download_stars(st.data());
		stars_host_.assign(st.begin(), st.end());
		return true;
	}

	/* ccf.cuh:1102-1258 */
	bool write_files(int verbose)
	{
		if (ccs_.empty())
		{
			std::cerr << "Error. run() must succeed before save()\n";
			return false;
		}
		std::string fname = outfile_prefix + "ccf_parameter_info.txt";
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
		out << "num_stars " << num_stars << "\n";
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
		out << "num_roots " << num_roots_ << "\n";
		out << "num_phi " << num_phi << "\n";
		out << "num_branches " << num_branches << "\n";
		out << "max_error_1/mu " << max_error_ << "\n";
		out << "t_init_roots " << t_init_roots_ << "\n";
		out << "t_ccs " << t_ccs_ << "\n";
		out << "t_caustics " << t_caustics_ << "\n";
		out.close();
		MLENS_VERBOSE("Done writing parameter info to file " << fname << "\n", verbose, 1);

		if (write_stars)
		{
			fname = outfile_prefix + "ccf_stars" + outfile_type;
			if (!write_star_file<T>(fname, num_stars, rectangular, corner_, theta_star, stars_host_.data()))
			{
				std::cerr << "Error. Unable to write star info to file " << fname << "\n";
				return false;
			}
		}
		const int nrows = num_roots_ * num_branches, ncols = num_phi / num_branches + 1;
		if (write_critical_curves)
		{
			fname = outfile_prefix + "ccf_ccs" + outfile_type;
			if (!write_array<Complex<T>>(fname, reinterpret_cast<const Complex<T>*>(ccs_.data()), nrows, ncols))
			{
				std::cerr << "Error. Unable to write ccs info to file " << fname << "\n";
				return false;
			}
		}
		if (write_caustics)
		{
			fname = outfile_prefix + "ccf_caustics" + outfile_type;
			if (!write_array<Complex<T>>(fname, reinterpret_cast<const Complex<T>*>(caustics_.data()), nrows, ncols))
			{
				std::cerr << "Error. Unable to write caustic info to file " << fname << "\n";
				return false;
			}
		}
		if (write_mu_length_scales)
		{
			fname = outfile_prefix + "ccf_mu_length_scales" + outfile_type;
			if (!write_array<T>(fname, mu_.data(), nrows, ncols))
			{
				std::cerr << "Error. Unable to write magnification length scales info to file " << fname << "\n";
				return false;
			}
		}
		return true;
	}
};

}  // namespace mlens
