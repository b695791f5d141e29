// lib_irs.so: a C ABI for IRS modelled on the reference's IPM ABI
// (src/lib_ipm.cu:12-83). The reference builds IRS only as a command-line
// program (src/main_ipm_irs.cu with -D is_irs, irs.cuh:1224-1247); these entry
// points expose the same IRS<T> fields and results to a library caller.
#include "../../include/lib_irs.h"

#include "map_maker.hpp"

struct IRSHandle
{
	mlens::MapMaker<irs_dtype, false> m;
};

#define HANDLE IRSHandle
#define DTYPE irs_dtype
#include "abi_common.hpp"

extern "C" {

IRSHandle* init() { return new IRSHandle(); }
void irs_free(IRSHandle* self) { delete self; }

MLENS_COMMON_FIELDS
MLENS_COMMON_ENTRIES

int* get_pixels(IRSHandle* self) { return self->m.get_ipixels(); }
int* get_pixels_minima(IRSHandle* self) { return self->m.get_ipixels_minima(); }
int* get_pixels_saddles(IRSHandle* self) { return self->m.get_ipixels_saddles(); }
double get_t_shoot_rays(IRSHandle* self) { return self->m.get_t_shoot(); }
long long get_num_rays_shot(IRSHandle* self) { return self->m.get_num_rays_shot(); }
long long get_num_rays_received(IRSHandle* self) { return self->m.get_num_rays_received(); }

bool irs_prepare(IRSHandle* self, int verbose) { MLENS_GUARD(return self->m.prepare(verbose), return false) }
bool irs_derive(IRSHandle* self, int verbose) { MLENS_GUARD(return self->m.derive(verbose), return false) }
bool irs_build(IRSHandle* self, int verbose) { MLENS_GUARD(return self->m.build(verbose), return false) }
bool irs_shoot(IRSHandle* self, int verbose, long long row0, long long row1)
{
	MLENS_GUARD(return self->m.shoot(verbose, row0, row1), return false)
}
bool irs_finish(IRSHandle* self, int verbose) { MLENS_GUARD(return self->m.finish(verbose), return false) }

bool irs_ray_positions(IRSHandle* self, long long row, double* out)
{
	if (!self->m.engine()) return false;
	MLENS_GUARD(self->m.engine()->irs_positions(self->m.lens(), self->m.grid(), mlens::Precision(self->m.precision), row, out);
				return true, return false)
}

}  // extern "C"
