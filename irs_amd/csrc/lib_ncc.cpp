// lib_ncc.so: the reference's NCC C ABI (src/lib_ncc.cu:12-46) over the MI355X
// caustic-crossing counter, with double values as the reference builds it.
#include "../../include/lib_ncc.h"

#include "ncc_maker.hpp"

struct NCCHandle
{
	mlens::NCCMaker<ncc_dtype> m;
};

#define HANDLE NCCHandle
#define DTYPE ncc_dtype
#include "abi_common.hpp"

extern "C" {

NCCHandle* init() { return new NCCHandle(); }
void ncc_free(NCCHandle* self) { delete self; }

MLENS_STRING_FIELD(infile_prefix, infile_prefix)
MLENS_FIELD(center_y1, DTYPE, center_y.re)
MLENS_FIELD(center_y2, DTYPE, center_y.im)
MLENS_FIELD(half_length_y1, DTYPE, half_length_y.re)
MLENS_FIELD(half_length_y2, DTYPE, half_length_y.im)
MLENS_FIELD(num_pixels_y1, int, num_pixels_y.re)
MLENS_FIELD(num_pixels_y2, int, num_pixels_y.im)
MLENS_FIELD(over_sample, int, over_sample)
MLENS_FIELD(write_maps, int, write_maps)
MLENS_FIELD(write_histograms, int, write_histograms)
MLENS_STRING_FIELD(outfile_prefix, outfile_prefix)

int* get_num_crossings(NCCHandle* self) { return self->m.get_num_crossings(); }
double get_t_ncc(NCCHandle* self) { return self->m.get_t_ncc(); }
double get_t_reduce(NCCHandle* self) { return self->m.get_t_reduce(); }

bool run(NCCHandle* self, int verbose) { MLENS_GUARD(return self->m.run(verbose), return false) }
bool save(NCCHandle* self, int verbose) { MLENS_GUARD(return self->m.save(verbose), return false) }

}  // extern "C"
