// File formats of the reference map makers, restated:
//   star field .bin  -- star.cuh:175-320 (read), 498-534 (write):
//                       int32 N, int32 rectangular, Complex<T> corner, T theta,
//                       then N x {T x1, T x2, T m}; float or double files are
//                       both accepted on read, recognised by their size
//   star field .txt  -- star.cuh:334-434: "x1 x2 m" lines, theta = 1, shape
//                       from the moment of inertia
//   map .bin         -- array_functions.cuh:280-306: int32 nrows, int32 ncols,
//                       then nrows*ncols raw T
//   histogram .txt   -- array_functions.cuh:318-349: "value count" lines for
//                       the non-zero bins
#pragma once

#include "common.hpp"

#include <algorithm>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

namespace mlens
{

struct StarFileData
{
	int num_stars = 0;
	int rectangular = 0;
	double corner_re = 0, corner_im = 0;
	double theta = 1;
	std::vector<double> stars;   // AoS x1, x2, m
};

template <typename F>
static bool read_bin_stars_as(std::ifstream& in, int n, StarFileData& d)
{
	F c[2];
	F th;
	in.read(reinterpret_cast<char*>(c), sizeof(c));
	in.read(reinterpret_cast<char*>(&th), sizeof(th));
	std::vector<F> buf(size_t(n) * 3);
	in.read(reinterpret_cast<char*>(buf.data()), buf.size() * sizeof(F));
	if (!in) return false;
	d.corner_re = c[0];
	d.corner_im = c[1];
	d.theta = th;
	d.stars.assign(buf.begin(), buf.end());
	return true;
}

inline bool read_star_file(const std::string& fname, StarFileData& d)
{
	std::filesystem::path path = fname;
	if (path.extension() == ".bin")
	{
		std::error_code err;
		std::uintmax_t fsize = std::filesystem::file_size(fname, err);
		if (err)
		{
			std::cerr << "Error determining size of star input file " << fname << "\n";
			return false;
		}
		std::ifstream in(fname, std::ios_base::binary);
		if (!in.is_open())
		{
			std::cerr << "Error. Failed to open file " << fname << "\n";
			return false;
		}
		int n = 0;
		in.read(reinterpret_cast<char*>(&n), sizeof(int));
		if (!in || n < 1)
		{
			std::cerr << "Error. Invalid num_stars input. num_stars must be an integer > 0\n";
			return false;
		}
		int rect = -1;
		in.read(reinterpret_cast<char*>(&rect), sizeof(int));
		if (rect != 0 && rect != 1)
		{
			std::cerr << "Error. Invalid rectangular input. rectangular must be 1 (rectangular) or 0 (circular).\n";
			return false;
		}
		d.num_stars = n;
		d.rectangular = rect;
		const std::uintmax_t head = 2 * sizeof(int);
		bool ok = false;
		if (fsize == head + 3 * sizeof(float) + std::uintmax_t(n) * 3 * sizeof(float))
			ok = read_bin_stars_as<float>(in, n, d);
		else if (fsize == head + 3 * sizeof(double) + std::uintmax_t(n) * 3 * sizeof(double))
			ok = read_bin_stars_as<double>(in, n, d);
		else
		{
			std::cerr << "Error. Star input file " << fname
					  << " does not contain validly formatted single or double precision stars and accompanying information.\n";
			return false;
		}
		if (!ok)
		{
			std::cerr << "Error. Failed to read star input file " << fname << "\n";
			return false;
		}
		if (d.corner_re < 0 || d.corner_im < 0)
		{
			std::cerr << "Error. Real and imaginary parts of the corner of the star field must both be >= 0\n";
			return false;
		}
		if (d.theta < std::numeric_limits<double>::min())
		{
			std::cerr << "Error. theta_star must be >= " << std::numeric_limits<double>::min() << "\n";
			return false;
		}
		return true;
	}
	if (path.extension() == ".txt")
	{
		std::ifstream in(fname);
		if (!in.is_open())
		{
			std::cerr << "Error. Failed to open file " << fname << "\n";
			return false;
		}
		double x1, x2, m;
		d.stars.clear();
		while (in >> x1 >> x2 >> m)
		{
			d.stars.push_back(x1);
			d.stars.push_back(x2);
			d.stars.push_back(m);
		}
		d.num_stars = int(d.stars.size() / 3);
		if (d.num_stars < 1)
		{
			std::cerr << "Error. Invalid num_stars input. num_stars must be an integer > 0\n";
			return false;
		}
		double mx1 = 0, mx2 = 0, mrad = 0, mt = 0, inertia = 0;
		for (int i = 0; i < d.num_stars; i++)
		{
			double a = d.stars[3 * i], b = d.stars[3 * i + 1], mm = d.stars[3 * i + 2];
			mt += mm;
			inertia += mm * (a * a + b * b);
			mx1 = std::max(mx1, std::abs(a));
			mx2 = std::max(mx2, std::abs(b));
			mrad = std::max(mrad, std::hypot(a, b));
		}
		double icirc = mt * mrad * mrad / 2;
		double irect = mt * (mx1 * mx1 + mx2 * mx2) / 3;
		if (inertia / icirc > inertia / irect)
		{
			d.rectangular = 0;
			d.corner_re = mrad;
			d.corner_im = 0;
		}
		else
		{
			d.rectangular = 1;
			d.corner_re = mx1;
			d.corner_im = mx2;
		}
		d.theta = 1;
		return true;
	}
	std::cerr << "Error. Star input file " << fname << " is not a .bin or .txt file.\n";
	return false;
}

template <typename T>
inline bool write_star_file(const std::string& fname, int n, int rectangular, Complex<T> corner, T theta, const T* stars_aos)
{
	if (std::filesystem::path(fname).extension() != ".bin")
	{
		std::cerr << "Error. Star file " << fname << " is not a .bin file.\n";
		return false;
	}
	std::ofstream out(fname, std::ios_base::binary);
	if (!out.is_open())
	{
		std::cerr << "Error. Failed to open file " << fname << "\n";
		return false;
	}
	Complex<T> c = rectangular ? corner : Complex<T>(corner.abs(), 0);
	T cc[2] = {c.re, c.im};
	out.write(reinterpret_cast<const char*>(&n), sizeof(int));
	out.write(reinterpret_cast<const char*>(&rectangular), sizeof(int));
	out.write(reinterpret_cast<const char*>(cc), sizeof(cc));
	out.write(reinterpret_cast<const char*>(&theta), sizeof(T));
	out.write(reinterpret_cast<const char*>(stars_aos), std::streamsize(size_t(n) * 3 * sizeof(T)));
	return bool(out);
}

template <typename V>
inline bool write_array(const std::string& fname, const V* vals, int nrows, int ncols)
{
	if (std::filesystem::path(fname).extension() != ".bin")
	{
		std::cerr << "Error. File " << fname << " is not a .bin file.\n";
		return false;
	}
	std::ofstream out(fname, std::ios_base::binary);
	if (!out.is_open())
	{
		std::cerr << "Error. Failed to open file " << fname << "\n";
		return false;
	}
	out.write(reinterpret_cast<const char*>(&nrows), sizeof(int));
	out.write(reinterpret_cast<const char*>(&ncols), sizeof(int));
	out.write(reinterpret_cast<const char*>(vals), std::streamsize(size_t(nrows) * ncols * sizeof(V)));
	return bool(out);
}

// array_functions.cuh:173-267: int32 nrows, int32 ncols, then nrows*ncols
// complex values, double or float (told apart by the file size); returned as
// interleaved (re, im) doubles
inline bool read_complex_array(const std::string& fname, std::vector<double>& vals, int& nrows, int& ncols)
{
	if (std::filesystem::path(fname).extension() != ".bin")
	{
		std::cerr << "Error. File " << fname << " is not a .bin file.\n";
		return false;
	}
	std::error_code err;
	const std::uintmax_t fsize = std::filesystem::file_size(fname, err);
	if (err)
	{
		std::cerr << "Error determining size of input file " << fname << "\n";
		return false;
	}
	std::ifstream in(fname, std::ios_base::binary);
	if (!in.is_open())
	{
		std::cerr << "Error. Failed to open file " << fname << "\n";
		return false;
	}
	nrows = 0;
	ncols = 0;
	in.read(reinterpret_cast<char*>(&nrows), sizeof(int));
	in.read(reinterpret_cast<char*>(&ncols), sizeof(int));
	if (!in || nrows < 1 || ncols < 1)
	{
		std::cerr << "Error. File " << fname << " does not contain valid values for num_rows and num_cols.\n";
		return false;
	}
	const std::uintmax_t n = std::uintmax_t(nrows) * std::uintmax_t(ncols);
	vals.resize(size_t(2 * n));
	if (fsize == 2 * sizeof(int) + n * 2 * sizeof(double))
	{
		in.read(reinterpret_cast<char*>(vals.data()), std::streamsize(2 * n * sizeof(double)));
	}
	else if (fsize == 2 * sizeof(int) + n * 2 * sizeof(float))
	{
		std::vector<float> tmp(size_t(2 * n));
		in.read(reinterpret_cast<char*>(tmp.data()), std::streamsize(2 * n * sizeof(float)));
		std::copy(tmp.begin(), tmp.end(), vals.begin());
	}
	else
	{
		std::cerr << "Error. Binary file does not contain valid single or double precision numbers.\n";
		return false;
	}
	if (!in)
	{
		std::cerr << "Error. Failed to read file " << fname << "\n";
		return false;
	}
	return true;
}

inline bool write_histogram(const std::string& fname, const std::vector<long long>& hist, long long minval)
{
	if (std::filesystem::path(fname).extension() != ".txt")
	{
		std::cerr << "Error. File " << fname << " is not a .txt file.\n";
		return false;
	}
	std::ofstream out(fname);
	if (!out.is_open())
	{
		std::cerr << "Error. Failed to open file " << fname << "\n";
		return false;
	}
	for (size_t i = 0; i < hist.size(); i++)
		if (hist[i] != 0) out << (long long)(i) + minval << " " << hist[i] << "\n";
	return bool(out);
}

}  // namespace mlens
