// Shared definitions for the MI355X microlensing map engine (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <string>

namespace mlens
{

// Tree constants of the reference quadtree (tree_node.cuh:14-25).
constexpr int MAX_NUM_STARS_DIRECT = 32;
constexpr int MAX_EXPANSION_ORDER = 31;
constexpr int MAX_NUM_NEIGHBORS = 8;
constexpr int MAX_NUM_M2L = 27;
constexpr int MAX_NUM_P2L = 5;
constexpr int MAX_TAYLOR_SMOOTH = 101;   // ipm.cuh:81

constexpr double PI = 3.14159265358979323846;

// Minimal complex value type used by host code (derived parameters, I/O).
template <typename T>
struct Complex
{
	T re = 0;
	T im = 0;

	__host__ __device__ Complex() {}
	__host__ __device__ Complex(T r, T i = 0) : re(r), im(i) {}
	template <typename U>
	__host__ __device__ explicit Complex(const Complex<U>& c) : re(static_cast<T>(c.re)), im(static_cast<T>(c.im)) {}

	__host__ __device__ T abs() const { return std::hypot(re, im); }
	__host__ __device__ T arg() const { return std::atan2(im, re); }
	__host__ __device__ Complex conj() const { return Complex(re, -im); }
	__host__ __device__ Complex operator+(Complex o) const { return Complex(re + o.re, im + o.im); }
	__host__ __device__ Complex operator-(Complex o) const { return Complex(re - o.re, im - o.im); }
	__host__ __device__ Complex operator*(Complex o) const { return Complex(re * o.re - im * o.im, re * o.im + im * o.re); }
	__host__ __device__ Complex operator*(T s) const { return Complex(re * s, im * s); }
	__host__ __device__ Complex operator/(T s) const { return Complex(re / s, im / s); }
};

// Lens parameters needed by every deflection evaluation.
struct LensParams
{
	double kappa;        // total convergence
	double gamma;        // external shear
	double theta;        // Einstein radius of a unit mass
	double kappastar;    // convergence in stars (the input value, as the reference passes it)
	int rectangular;
	int approx;
	int taylor_smooth;
	double corner_re, corner_im;
};

// Image-plane lattice and source-plane map geometry.
struct GridParams
{
	double center_x[2];
	double half_length_x[2];
	double ray_half_sep[2];
	long long num_cells[2];   // cells (IPM) or rays (IRS) along x1 and x2
	double center_y[2];
	double half_length_y[2];
	int num_pixels[2];
};

}  // namespace mlens

// prints `what` to stdout when verbose >= level
#define MLENS_VERBOSE(what, verbose, level) \
	do                                      \
	{                                       \
		if ((verbose) >= (level))           \
		{                                   \
			std::cout << what;              \
		}                                   \
	} while (false)

#define MLENS_HIP_CHECK(expr)                                                                          \
	do                                                                                                 \
	{                                                                                                  \
		hipError_t _e = (expr);                                                                        \
		if (_e != hipSuccess)                                                                          \
		{                                                                                              \
			throw mlens::HipError(std::string(#expr) + ": " + hipGetErrorString(_e), __FILE__, __LINE__); \
		}                                                                                              \
	} while (false)

namespace mlens
{
struct HipError
{
	std::string msg;
	std::string file;
	int line;
	HipError(std::string m, const char* f, int l) : msg(std::move(m)), file(f), line(l) {}
};
}  // namespace mlens
