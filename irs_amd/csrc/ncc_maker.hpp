// Host orchestration of the number of caustic crossings: the NCC<T> class of
// the reference (ncc.cuh:20-465) with the same public fields and defaults, the
// same validation and messages, and run() / save(). The caustics are read from
// <infile_prefix>ccf_caustics.bin, as the critical-curve finder writes them.
// As for the maps: validation comes before any device call, the device is not
// reset, the returned array is a host copy valid until this object's next
// run(), and the phases are timed with HIP events (seconds).
#pragma once

#include "common.hpp"
#include "io.hpp"
#include "ncc.hpp"

#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mlens
{

template <typename T>
class NCCMaker
{
public:
	/* public input fields and defaults (ncc.cuh:28-35) */
	std::string infile_prefix = "./";
	Complex<T> center_y = Complex<T>(0, 0);
	Complex<T> half_length_y = Complex<T>(5, 5);
	Complex<int> num_pixels_y = Complex<int>(1000, 1000);
	int over_sample = 2;
	int write_maps = 1;
	int write_histograms = 1;
	std::string outfile_prefix = "./";

	/* ncc.cuh:441-452 */
	bool run(int verbose)
	{
		try
		{
			if (!check_input_params(verbose)) return false;
			if (!read_caustics(verbose)) return false;
			if (!set_device()) return false;
			if (!calculate_num_caustic_crossings(verbose)) return false;
			return create_histograms(verbose);
		}
		catch (const HipError& e)
		{
			std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";
			return false;
		}
		catch (const std::exception& e)
		{
			std::cerr << "Error. " << e.what() << "\n";
			return false;
		}
	}

	bool save(int verbose) { return write_files(verbose); }

	int* get_num_crossings() { return num_.empty() ? nullptr : num_.data(); }
	double get_t_ncc() const { return t_ncc_; }
	double get_t_reduce() const { return t_reduce_; }
	long long get_num_crossings_recorded() const { return stats_.crossings; }

private:
	const std::string caustics_file = "ccf_caustics.bin";
	const std::string outfile_type = ".bin";

	std::unique_ptr<NCCDevice> device_;
	int num_rows_ = 0, num_cols_ = 0;
	std::vector<double> caustics_;
	std::vector<int> num_;
	std::vector<long long> histogram_;
	int min_num_ = 0, max_num_ = 0;
	double t_ncc_ = 0, t_reduce_ = 0;
	NCCStats stats_;

	/* ncc.cuh:150-192 */
	bool check_input_params(int verbose)
	{
		MLENS_VERBOSE("Checking input parameters...\n", verbose, 3);
		const T tmin = std::numeric_limits<T>::min();
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
		if (over_sample < 0)
		{
			std::cerr << "Error. over_sample must be an integer >= 0\n";
			return false;
		}
		if (write_maps != 0 && write_maps != 1)
		{
			std::cerr << "Error. write_maps must be 1 (true) or 0 (false).\n";
			return false;
		}
		if (write_histograms != 0 && write_histograms != 1)
		{
			std::cerr << "Error. write_histograms must be 1 (true) or 0 (false).\n";
			return false;
		}
		MLENS_VERBOSE("Done checking input parameters.\n\n", verbose, 3);
		return true;
	}

	/* ncc.cuh:194-222 */
	bool read_caustics(int verbose)
	{
		MLENS_VERBOSE("Reading in caustics...\n", verbose, 1);
		const std::string fname = infile_prefix + caustics_file;
		if (!read_complex_array(fname, caustics_, num_rows_, num_cols_))
		{
			std::cerr << "Error. Unable to read caustics from file " << fname << "\n";
			return false;
		}
		if (num_rows_ < 1 || num_cols_ < 2)
		{
			std::cerr << "Error. File " << fname << " does not contain valid values for num_rows and num_cols.\n";
			return false;
		}
		MLENS_VERBOSE("Done reading in caustics: " << num_rows_ << " rows of " << num_cols_ << " points.\n\n", verbose, 1);
		return true;
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
		if (!device_) device_ = std::make_unique<NCCDevice>();
		return true;
	}

	/* ncc.cuh:256-330 */
	bool calculate_num_caustic_crossings(int verbose)
	{
		MLENS_VERBOSE("Calculating number of caustic crossings...\n", verbose, 1);
		NCCGrid g;
		g.center_re = double(center_y.re);
		g.center_im = double(center_y.im);
		g.half_length_re = double(half_length_y.re);
		g.half_length_im = double(half_length_y.im);
		g.num_pixels_re = num_pixels_y.re;
		g.num_pixels_im = num_pixels_y.im;
		g.over_sample = over_sample;
		num_.assign(size_t(num_pixels_y.re) * size_t(num_pixels_y.im), 0);
		device_->count(caustics_.data(), num_rows_, num_cols_, g, num_.data(), stats_);
		t_ncc_ = stats_.t_ncc_ms / 1000;
		t_reduce_ = stats_.t_reduce_ms / 1000;
		if (stats_.dropped > 0)
			std::cerr << "Error. Caustic crossing takes place outside the desired region (" << stats_.dropped
					  << " crossings).\n";
		MLENS_VERBOSE("Done calculating number of caustic crossings. Elapsed time: " << t_ncc_ << " seconds.\n\n", verbose, 1);
		MLENS_VERBOSE("Downsampling number of caustic crossings... Elapsed time: " << t_reduce_ << " seconds.\n\n", verbose, 1);
		min_num_ = stats_.min;
		max_num_ = stats_.max;
		if (min_num_ < 0)
		{
			std::cerr << "Error. Number of caustic crossings should be >= 0\n";
			return false;
		}
		return true;
	}

	/* ncc.cuh:332-370 */
	bool create_histograms(int verbose)
	{
		histogram_.clear();
		if (!write_histograms) return true;
		MLENS_VERBOSE("Creating histograms...\n", verbose, 2);
		histogram_.assign(size_t(max_num_ - min_num_ + 1), 0);
		device_->histogram(min_num_, max_num_, histogram_.data());
		MLENS_VERBOSE("Done creating histograms.\n\n", verbose, 2);
		return true;
	}

	/* ncc.cuh:372-438 */
	bool write_files(int verbose)
	{
		if (num_.empty())
		{
			std::cerr << "Error. There is no number of caustic crossings map to save; call run() first.\n";
			return false;
		}
		MLENS_VERBOSE("Writing parameter info...\n", verbose, 2);
		std::string fname = outfile_prefix + "ncc_parameter_info.txt";
		std::ofstream out(fname);
		if (!out.is_open())
		{
			std::cerr << "Error. Failed to open file " << fname << "\n";
			return false;
		}
		out.precision(9);
		out << "center_y1 " << center_y.re << "\n";
		out << "center_y2 " << center_y.im << "\n";
		out << "half_length_y1 " << half_length_y.re << "\n";
		out << "half_length_y2 " << half_length_y.im << "\n";
		out << "num_pixels_y1 " << num_pixels_y.re << "\n";
		out << "num_pixels_y2 " << num_pixels_y.im << "\n";
		out << "over_sample " << over_sample << "\n";
		out << "t_ncc " << t_ncc_ << "\n";
		out << "t_reduce " << t_reduce_ << "\n";
		out.close();
		MLENS_VERBOSE("Done writing parameter info to file " << fname << "\n", verbose, 1);
		if (write_histograms)
		{
			fname = outfile_prefix + "ncc_ncc_numpixels.txt";
			if (!write_histogram(fname, histogram_, min_num_))
			{
				std::cerr << "Error. Unable to write caustic crossings histogram to file " << fname << "\n";
				return false;
			}
			MLENS_VERBOSE("Done writing number of caustic crossings histogram to file " << fname << "\n", verbose, 1);
		}
		if (write_maps)
		{
			fname = outfile_prefix + "ncc_ncc" + outfile_type;
			if (!write_array<int>(fname, num_.data(), num_pixels_y.im, num_pixels_y.re))
			{
				std::cerr << "Error. Unable to write number of caustic crossings to file " << fname << "\n";
				return false;
			}
			MLENS_VERBOSE("Done writing number of caustic crossings to file " << fname << "\n", verbose, 1);
		}
		return true;
	}
};

}  // namespace mlens
