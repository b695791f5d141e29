/*
 * lib_irs.h -- C ABI of lib_irs.so (float API, mixed-precision engine by default).
 *
 * The reference builds inverse ray shooting only as a command-line program
 * (src/main_ipm_irs.cu:9-36 with -D is_irs; class IRS<T> in include/irs.cuh).
 * It has no IRS library, so there is no reference binding to replace; this ABI
 * mirrors the IPM one (src/lib_ipm.cu:12-83) field for field, with the IRS
 * defaults (light_loss 0.01, num_rays_y 100, irs.cuh:51,59) and integer ray
 * counts as the map (irs_functions.cuh:111-212).
 */
#ifndef MLENS_LIB_IRS_H
#define MLENS_LIB_IRS_H

#include <stdbool.h>

#ifndef IRS_DTYPE
#define IRS_DTYPE float
#endif
typedef IRS_DTYPE irs_dtype;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct IRSHandle IRSHandle;

IRSHandle* init(void);
void irs_free(IRSHandle* self);

/* setters and getters with the IPM names and meaning (lib_ipm.cu:17-69) */
void set_kappa_tot(IRSHandle* self, irs_dtype val);
void set_shear(IRSHandle* self, irs_dtype val);
void set_kappa_star(IRSHandle* self, irs_dtype val);
void set_theta_star(IRSHandle* self, irs_dtype val);
void set_mass_function(IRSHandle* self, const char* val);
void set_m_solar(IRSHandle* self, irs_dtype val);
void set_m_lower(IRSHandle* self, irs_dtype val);
void set_m_upper(IRSHandle* self, irs_dtype val);
void set_light_loss(IRSHandle* self, irs_dtype val);
void set_rectangular(IRSHandle* self, int val);
void set_approx(IRSHandle* self, int val);
void set_safety_scale(IRSHandle* self, irs_dtype val);
void set_starfile(IRSHandle* self, const char* val);
void set_center_y1(IRSHandle* self, irs_dtype val);
void set_center_y2(IRSHandle* self, irs_dtype val);
void set_half_length_y1(IRSHandle* self, irs_dtype val);
void set_half_length_y2(IRSHandle* self, irs_dtype val);
void set_num_pixels_y1(IRSHandle* self, int val);
void set_num_pixels_y2(IRSHandle* self, int val);
void set_num_rays_y(IRSHandle* self, int val);
void set_random_seed(IRSHandle* self, int val);
void set_write_stars(IRSHandle* self, int val);
void set_write_maps(IRSHandle* self, int val);
void set_write_parities(IRSHandle* self, int val);
void set_write_histograms(IRSHandle* self, int val);
void set_outfile_prefix(IRSHandle* self, const char* val);
irs_dtype get_kappa_tot(IRSHandle* self);
irs_dtype get_shear(IRSHandle* self);
irs_dtype get_kappa_star(IRSHandle* self);
irs_dtype get_theta_star(IRSHandle* self);
const char* get_mass_function(IRSHandle* self);
irs_dtype get_m_solar(IRSHandle* self);
irs_dtype get_m_lower(IRSHandle* self);
irs_dtype get_m_upper(IRSHandle* self);
irs_dtype get_light_loss(IRSHandle* self);
int get_rectangular(IRSHandle* self);
int get_approx(IRSHandle* self);
irs_dtype get_safety_scale(IRSHandle* self);
const char* get_starfile(IRSHandle* self);
irs_dtype get_center_y1(IRSHandle* self);
irs_dtype get_center_y2(IRSHandle* self);
irs_dtype get_half_length_y1(IRSHandle* self);
irs_dtype get_half_length_y2(IRSHandle* self);
int get_num_pixels_y1(IRSHandle* self);
int get_num_pixels_y2(IRSHandle* self);
int get_num_rays_y(IRSHandle* self);
int get_random_seed(IRSHandle* self);
int get_write_stars(IRSHandle* self);
int get_write_maps(IRSHandle* self);
int get_write_parities(IRSHandle* self);
int get_write_histograms(IRSHandle* self);
const char* get_outfile_prefix(IRSHandle* self);

int get_num_stars(IRSHandle* self);
irs_dtype get_corner_x1(IRSHandle* self);
irs_dtype get_corner_x2(IRSHandle* self);
irs_dtype* get_stars(IRSHandle* self);
int* get_pixels(IRSHandle* self);            /* ray counts, row-major n2 x n1, row 0 = top */
int* get_pixels_minima(IRSHandle* self);
int* get_pixels_saddles(IRSHandle* self);
double get_t_shoot_rays(IRSHandle* self);     /* seconds (irs.cuh:927) */
long long get_num_rays_shot(IRSHandle* self);      /* irs.cuh:930-932 */
long long get_num_rays_received(IRSHandle* self);  /* irs.cuh:934-936 */

bool run(IRSHandle* self, int verbose);
bool save(IRSHandle* self, int verbose);

/* staged run, precision, tree inspection: as in lib_ipm.h */
void set_expansion_order(IRSHandle* self, int val);
int get_expansion_order(IRSHandle* self);
void set_precision(IRSHandle* self, int val);
int get_precision(IRSHandle* self);
/* M2L of the FMM (fmm.cuh:281-374): 0 FP64 recurrences; 1 batched GEMMs on
   MFMA, FP32 for p <= 20 and FP64 above; 2 FP32 MFMA; 3 FP64 MFMA (default) */
void set_m2l_mode(IRSHandle* self, int val);
int get_m2l_mode(IRSHandle* self);
/* lattice rows per tile of the shooting kernel: 0 (default) chooses them per
   launch, a positive value (at most 65536) forces that height */
void set_tile_rows(IRSHandle* self, int val);
int get_tile_rows(IRSHandle* self);
int get_last_tile_rows(IRSHandle* self);                  /* the height the last shoot used */
double get_t_m2l(IRSHandle* self);          /* seconds in the M2L kernels of the last build (MFMA modes) */
long long get_m2l_pairs(IRSHandle* self);   /* M2L (target, source) pairs of the last build (MFMA modes) */
int get_m2l_fp32(IRSHandle* self);          /* 1 if the last build's M2L ran on FP32 MFMA */
int get_tree_levels(IRSHandle* self);
irs_dtype get_alpha_error(IRSHandle* self);
int get_num_ray_threads_1(IRSHandle* self);
int get_num_ray_threads_2(IRSHandle* self);
irs_dtype get_root_half_length(IRSHandle* self);
double get_mean_near_stars(IRSHandle* self);
int get_taylor_smooth(IRSHandle* self);
irs_dtype get_center_x1(IRSHandle* self);                  /* derived image-plane geometry (ipm.cuh:460-470) */
irs_dtype get_center_x2(IRSHandle* self);
irs_dtype get_half_length_x1(IRSHandle* self);
irs_dtype get_half_length_x2(IRSHandle* self);
irs_dtype get_ray_half_sep_1(IRSHandle* self);
irs_dtype get_ray_half_sep_2(IRSHandle* self);
irs_dtype get_kappa_star_actual(IRSHandle* self);
irs_dtype get_mean_mass(IRSHandle* self);
irs_dtype get_mean_mass2(IRSHandle* self);
irs_dtype get_mean_mass_actual(IRSHandle* self);
irs_dtype get_mean_mass2_actual(IRSHandle* self);            /* star moments of the realised field (star.cuh:124-161) */
irs_dtype get_mean_mass2_ln_mass(IRSHandle* self);
irs_dtype get_mean_mass2_ln_mass_actual(IRSHandle* self);
irs_dtype get_m_lower_actual(IRSHandle* self);
irs_dtype get_m_upper_actual(IRSHandle* self);
irs_dtype get_mu_ave(IRSHandle* self);
int get_num_nodes(IRSHandle* self);
long long get_cells_shot(IRSHandle* self);
bool irs_prepare(IRSHandle* self, int verbose);
/* validation and derived geometry only (no device work), as ipm_derive */
bool irs_derive(IRSHandle* self, int verbose);
bool irs_build(IRSHandle* self, int verbose);
bool irs_shoot(IRSHandle* self, int verbose, long long row0, long long row1);
bool irs_finish(IRSHandle* self, int verbose);
/* one lattice row of rays: out[3 i] = w1, out[3 i + 1] = w2 (source plane,
   before subtracting center_y), out[3 i + 2] = the pixel index the kernel bins
   the ray into (-1 outside the map, -2 NaN); for bit-exact binning tests */
bool irs_ray_positions(IRSHandle* self, long long row, double* out);

#ifdef __cplusplus
}
#endif

#endif
