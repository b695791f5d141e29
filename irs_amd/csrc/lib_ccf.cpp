// lib_ccf.so: the reference's CCF C ABI (src/lib_ccf.cu:14-70) over the MI355X
// critical-curve solver, with double values as the reference builds it.
#include "../../include/lib_ccf.h"

#include "ccf_maker.hpp"

struct CCFHandle
{
	mlens::CCFMaker<ccf_dtype> m;
};

#define HANDLE CCFHandle
#define DTYPE ccf_dtype
#include "abi_common.hpp"

extern "C" {

CCFHandle* init() { return new CCFHandle(); }
void ccf_free(CCFHandle* self) { delete self; }

MLENS_FIELD(kappa_tot, DTYPE, kappa_tot)
MLENS_FIELD(shear, DTYPE, shear)
MLENS_FIELD(kappa_star, DTYPE, kappa_star)
MLENS_FIELD(theta_star, DTYPE, theta_star)
MLENS_STRING_FIELD(mass_function, mass_function_str)
MLENS_FIELD(m_solar, DTYPE, m_solar)
MLENS_FIELD(m_lower, DTYPE, m_lower)
MLENS_FIELD(m_upper, DTYPE, m_upper)
MLENS_FIELD(rectangular, int, rectangular)
MLENS_FIELD(approx, int, approx)
MLENS_FIELD(safety_scale, DTYPE, safety_scale)
MLENS_FIELD(num_stars, int, num_stars)
MLENS_STRING_FIELD(starfile, starfile)
MLENS_FIELD(num_phi, int, num_phi)
MLENS_FIELD(num_branches, int, num_branches)
MLENS_FIELD(random_seed, int, random_seed)
MLENS_FIELD(write_stars, int, write_stars)
MLENS_FIELD(write_critical_curves, int, write_critical_curves)
MLENS_FIELD(write_caustics, int, write_caustics)
MLENS_FIELD(write_mu_length_scales, int, write_mu_length_scales)
MLENS_STRING_FIELD(outfile_prefix, outfile_prefix)

int get_num_roots(CCFHandle* self) { return self->m.get_num_roots(); }
ccf_dtype get_corner_x1(CCFHandle* self) { return self->m.get_corner().re; }
ccf_dtype get_corner_x2(CCFHandle* self) { return self->m.get_corner().im; }
ccf_dtype* get_stars(CCFHandle* self) { return self->m.get_stars(); }
ccf_dtype* get_critical_curves(CCFHandle* self) { return self->m.get_critical_curves(); }
ccf_dtype* get_caustics(CCFHandle* self) { return self->m.get_caustics(); }
ccf_dtype* get_mu_length_scales(CCFHandle* self) { return self->m.get_mu_length_scales(); }
double get_t_ccs(CCFHandle* self) { return self->m.get_t_ccs(); }
double get_t_init_roots(CCFHandle* self) { return self->m.get_t_init_roots(); }
double get_max_error(CCFHandle* self) { return self->m.get_max_error(); }
int get_expansion_order(CCFHandle* self) { return self->m.get_expansion_order(); }
int get_tree_levels(CCFHandle* self) { return self->m.get_tree_levels(); }

bool run(CCFHandle* self, int verbose) { MLENS_GUARD(return self->m.run(verbose), return false) }
bool save(CCFHandle* self, int verbose) { MLENS_GUARD(return self->m.save(verbose), return false) }

}  // extern "C"
