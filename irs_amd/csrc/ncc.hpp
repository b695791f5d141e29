// Number of caustic crossings on the GPU (ncc_functions.cuh, ncc.cuh of
// weisluke/irs): for every pixel of a source-plane grid, the number of
// caustic curves that enclose the pixel's centre, counted by the crossings of
// a ray from the centre to +infinity in y1, on a grid over-sampled by
// 2^over_sample per axis and reduced by the maximum over each block.
#pragma once

#include "engine.hpp"

#include <vector>

namespace mlens
{

struct NCCGrid
{
	double center_re = 0, center_im = 0;            // centre of the map
	double half_length_re = 5, half_length_im = 5;  // half-lengths of the map
	int num_pixels_re = 1000, num_pixels_im = 1000; // final pixels per axis
	int over_sample = 2;                            // log2 of the sub-pixels per pixel and axis
};

struct NCCStats
{
	double t_ncc_ms = 0;      // the crossings (HIP events)
	double t_reduce_ms = 0;   // row sums and block maxima onto the final grid (HIP events)
	long long crossings = 0;  // crossings recorded over the fine grid
	long long dropped = 0;    // crossings in rows outside the grid (the reference reports each)
	int min = 0, max = 0;     // extremes of the final map
};

class NCCDevice
{
public:
	NCCDevice();
	~NCCDevice();
	NCCDevice(const NCCDevice&) = delete;
	NCCDevice& operator=(const NCCDevice&) = delete;

	// caustics: nrows x ncols complex points (re, im interleaved, row-major, host memory);
	// out: num_pixels_im x num_pixels_re counts, row 0 at the top (host memory)
	void count(const double* caustics, int nrows, int ncols, const NCCGrid& g, int* out, NCCStats& st);
	// hist[v - vmin] = pixels of the last map with value v, for v in [vmin, vmax]
	void histogram(int vmin, int vmax, long long* hist);

private:
	hipStream_t stream_ = nullptr;
	hipEvent_t ev0_ = nullptr, ev1_ = nullptr, ev2_ = nullptr;
	DeviceBuffer<double> caustics_;
	DeviceBuffer<int> fine_, map_, extremes_, row_ext_;
	DeviceBuffer<unsigned long long> counters_, hist_;
	long long map_pixels_ = 0;
};

}  // namespace mlens
