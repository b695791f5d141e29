// lib_ipm.so / lib_ipm_double.so: the reference's IPM C ABI (src/lib_ipm.cu:12-83)
// over the MI355X engine. Built twice, with IPM_DTYPE=float and IPM_DTYPE=double.
#include "../../include/lib_ipm.h"

#include "map_maker.hpp"

struct IPMHandle
{
	mlens::MapMaker<ipm_dtype, true> m;
};

#define HANDLE IPMHandle
#define DTYPE ipm_dtype
#include "abi_common.hpp"

extern "C" {

IPMHandle* init() { return new IPMHandle(); }
void ipm_free(IPMHandle* self) { delete self; }

MLENS_COMMON_FIELDS
MLENS_COMMON_ENTRIES

void set_kernel(IPMHandle* self, int val) { self->m.kernel = val; }
int get_kernel(IPMHandle* self) { return self->m.kernel; }

ipm_dtype* get_pixels(IPMHandle* self) { return self->m.get_pixels(); }
ipm_dtype* get_pixels_minima(IPMHandle* self) { return self->m.get_pixels_minima(); }
ipm_dtype* get_pixels_saddles(IPMHandle* self) { return self->m.get_pixels_saddles(); }
double get_t_shoot_cells(IPMHandle* self) { return self->m.get_t_shoot(); }

bool ipm_prepare(IPMHandle* self, int verbose) { MLENS_GUARD(return self->m.prepare(verbose), return false) }
bool ipm_build(IPMHandle* self, int verbose) { MLENS_GUARD(return self->m.build(verbose), return false) }
bool ipm_shoot(IPMHandle* self, int verbose, long long row0, long long row1)
{
	MLENS_GUARD(return self->m.shoot(verbose, row0, row1), return false)
}
bool ipm_finish(IPMHandle* self, int verbose) { MLENS_GUARD(return self->m.finish(verbose), return false) }

bool ipm_deflect(IPMHandle* self, const double* z, int n, double* w, double* mu)
{
	if (!self->m.engine()) return false;
	MLENS_GUARD(self->m.engine()->deflect_points(self->m.lens(), z, n, mlens::Precision(self->m.precision), w, mu); return true,
		return false)
}

bool ipm_export_tree(IPMHandle* self, double* geom, int* info, int* lists)
{
	if (!self->m.engine()) return false;
	MLENS_GUARD(self->m.engine()->export_tree(geom, info, lists); return true, return false)
}

bool ipm_export_coeffs(IPMHandle* self, int which, double* out)
{
	if (!self->m.engine()) return false;
	MLENS_GUARD(self->m.engine()->export_coeffs(which, out); return true, return false)
}

}  // extern "C"
