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
bool ipm_derive(IPMHandle* self, int verbose) { MLENS_GUARD(return self->m.derive(verbose), return false) }
bool ipm_build(IPMHandle* self, int verbose) { MLENS_GUARD(return self->m.build(verbose), return false) }
bool ipm_shoot(IPMHandle* self, int verbose, long long row0, long long row1)
{
	MLENS_GUARD(return self->m.shoot(verbose, row0, row1), return false)
}
bool ipm_finish(IPMHandle* self, int verbose) { MLENS_GUARD(return self->m.finish(verbose), return false) }

bool ipm_deflect(IPMHandle* self, const double* z, int n, double* w, double* mu)
{
	if (!self->m.engine() || !self->m.is_built())
	{
		std::cerr << "Error. build() must succeed before ipm_deflect()\n";
		return false;
	}
	MLENS_GUARD(self->m.engine()->deflect_points(self->m.lens(), z, n, mlens::Precision(self->m.precision), w, mu); return true,
		return false)
}

bool ipm_deflect_noalloc(IPMHandle* self, const double* z, int n, double* w, double* mu)
{
	if (!self->m.engine() || !self->m.is_built())
	{
		std::cerr << "Error. build() must succeed before ipm_deflect_noalloc()\n";
		return false;
	}
	MLENS_GUARD(self->m.engine()->deflect_points_noalloc(self->m.lens(), z, n, mlens::Precision(self->m.precision), w, mu);
				return true, return false)
}

long long ipm_window_cells(IPMHandle* self, int x0, int y0, int width, int height, double margin, long long* records,
	long long capacity)
{
	if (!self->m.engine() || !self->m.is_built())
	{
		std::cerr << "Error. build() must succeed before ipm_window_cells()\n";
		return -1;
	}
	MLENS_GUARD(return self->m.engine()->window_cells(self->m.lens(), self->m.grid(), x0, y0, width, height, margin,
					records, capacity),
		return -1)
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

/* many maps in one launch */
struct IPMBatch
{
	mlens::BatchShooter b;
	double t = 0;
};

IPMBatch* ipm_batch_create() { MLENS_GUARD(return new IPMBatch(), return nullptr) }
void ipm_batch_free(IPMBatch* self) { delete self; }
int ipm_batch_add(IPMBatch* self, IPMHandle* m)
{
	MLENS_GUARD(if (!m->m.prepare_batch_map()) return -1; return self->b.add(m->m.engine(), m->m.lens(), m->m.grid()),
		return -1)
}
bool ipm_batch_shoot(IPMBatch* self, int verbose) { MLENS_GUARD(self->t = self->b.shoot(verbose) / 1000.0; return true, return false) }
double get_batch_t_shoot(IPMBatch* self) { return self->t; }
long long get_batch_cells(IPMBatch* self) { return self->b.cells(); }
int get_batch_size(IPMBatch* self) { return self->b.size(); }
bool ipm_batch_histograms(IPMBatch* self, unsigned* out) { MLENS_GUARD(self->b.histograms(out); return true, return false) }
int get_batch_tile_h(IPMBatch* self) { return self->b.tile_h(); }
int get_batch_blocks_per_cu(IPMBatch* self) { return self->b.blocks_per_cu(); }
double get_batch_t_tune(IPMBatch* self) { return self->b.tune_ms() / 1000.0; }
bool ipm_batch_set_tuning(IPMBatch* self, int tile_h, int blocks_per_cu)
{
	MLENS_GUARD(self->b.set_tuning(tile_h, blocks_per_cu); return true, return false)
}

}  // extern "C"
