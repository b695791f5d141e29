/*
 * lib_ipm.h -- C ABI of lib_ipm.so (float) and lib_ipm_double.so (double).
 *
 * Drop-in replacement for the reference's src/lib_ipm.cu:12-83: the same 63
 * unprefixed extern "C" symbols with the same argument meaning, so the
 * reference's own ctypes binding (microlensing/IPM/lib_ipm.py:5-135 and
 * lib_ipm_double.py) loads it unchanged. Define IPM_DTYPE to double before
 * including this header to describe lib_ipm_double.so.
 *
 * The object returned by init() is opaque. Arrays returned by the getters are
 * host memory owned by the object, valid until that object's next run() or
 * ipm_free(). run() validates its parameters before it touches the GPU and
 * returns false with a message on stderr on any error.
 *
 * Entries marked "extension" have no counterpart in the reference.
 */
#ifndef MLENS_LIB_IPM_H
#define MLENS_LIB_IPM_H

#include <stdbool.h>

#ifndef IPM_DTYPE
#define IPM_DTYPE float
#endif
typedef IPM_DTYPE ipm_dtype;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct IPMHandle IPMHandle;

/* lib_ipm.cu:15 */
IPMHandle* init(void);

/* setters, lib_ipm.cu:17-42 */
void set_kappa_tot(IPMHandle* self, ipm_dtype val);
void set_shear(IPMHandle* self, ipm_dtype val);
void set_kappa_star(IPMHandle* self, ipm_dtype val);
void set_theta_star(IPMHandle* self, ipm_dtype val);
void set_mass_function(IPMHandle* self, const char* val);
void set_m_solar(IPMHandle* self, ipm_dtype val);
void set_m_lower(IPMHandle* self, ipm_dtype val);
void set_m_upper(IPMHandle* self, ipm_dtype val);
void set_light_loss(IPMHandle* self, ipm_dtype val);
void set_rectangular(IPMHandle* self, int val);
void set_approx(IPMHandle* self, int val);
void set_safety_scale(IPMHandle* self, ipm_dtype val);
void set_starfile(IPMHandle* self, const char* val);
void set_center_y1(IPMHandle* self, ipm_dtype val);
void set_center_y2(IPMHandle* self, ipm_dtype val);
void set_half_length_y1(IPMHandle* self, ipm_dtype val);
void set_half_length_y2(IPMHandle* self, ipm_dtype val);
void set_num_pixels_y1(IPMHandle* self, int val);
void set_num_pixels_y2(IPMHandle* self, int val);
void set_num_rays_y(IPMHandle* self, int val);
void set_random_seed(IPMHandle* self, int val);
void set_write_stars(IPMHandle* self, int val);
void set_write_maps(IPMHandle* self, int val);
void set_write_parities(IPMHandle* self, int val);
void set_write_histograms(IPMHandle* self, int val);
void set_outfile_prefix(IPMHandle* self, const char* val);

/* getters, lib_ipm.cu:44-69 */
ipm_dtype get_kappa_tot(IPMHandle* self);
ipm_dtype get_shear(IPMHandle* self);
ipm_dtype get_kappa_star(IPMHandle* self);
ipm_dtype get_theta_star(IPMHandle* self);
const char* get_mass_function(IPMHandle* self);
ipm_dtype get_m_solar(IPMHandle* self);
ipm_dtype get_m_lower(IPMHandle* self);
ipm_dtype get_m_upper(IPMHandle* self);
ipm_dtype get_light_loss(IPMHandle* self);
int get_rectangular(IPMHandle* self);
int get_approx(IPMHandle* self);
ipm_dtype get_safety_scale(IPMHandle* self);
const char* get_starfile(IPMHandle* self);
ipm_dtype get_center_y1(IPMHandle* self);
ipm_dtype get_center_y2(IPMHandle* self);
ipm_dtype get_half_length_y1(IPMHandle* self);
ipm_dtype get_half_length_y2(IPMHandle* self);
int get_num_pixels_y1(IPMHandle* self);
int get_num_pixels_y2(IPMHandle* self);
int get_num_rays_y(IPMHandle* self);
int get_random_seed(IPMHandle* self);
int get_write_stars(IPMHandle* self);
int get_write_maps(IPMHandle* self);
int get_write_parities(IPMHandle* self);
int get_write_histograms(IPMHandle* self);
const char* get_outfile_prefix(IPMHandle* self);

/* derived getters, lib_ipm.cu:71-78 */
int get_num_stars(IPMHandle* self);
ipm_dtype get_corner_x1(IPMHandle* self);
ipm_dtype get_corner_x2(IPMHandle* self);
ipm_dtype* get_stars(IPMHandle* self);           /* AoS [x1, x2, m] * num_stars, tree order */
ipm_dtype* get_pixels(IPMHandle* self);          /* row-major num_pixels_y2 x num_pixels_y1, row 0 = top */
ipm_dtype* get_pixels_minima(IPMHandle* self);
ipm_dtype* get_pixels_saddles(IPMHandle* self);
double get_t_shoot_cells(IPMHandle* self);      /* seconds */

/* actions, lib_ipm.cu:80-81 */
bool run(IPMHandle* self, int verbose);
bool save(IPMHandle* self, int verbose);

/* ---- extensions ---- */
void ipm_free(IPMHandle* self);                       /* destroys the object (the reference has none) */
void set_expansion_order(IPMHandle* self, int val);   /* 0: automatic order (ipm.cuh:848-860); else forced, <= 31 */
int get_expansion_order(IPMHandle* self);
void set_precision(IPMHandle* self, int val);         /* 0: FP64 throughout; 1: mixed FP32/FP64 (float library default) */
int get_precision(IPMHandle* self);
/* M2L of the FMM (fmm.cuh:281-374): 0 FP64 recurrences; 1 batched GEMMs on
   MFMA, FP32 for p <= 20 and FP64 above; 2 FP32 MFMA; 3 FP64 MFMA (default) */
void set_m2l_mode(IPMHandle* self, int val);
int get_m2l_mode(IPMHandle* self);
/* lattice rows per tile of the shooting kernel: 0 (default) chooses them per
   launch, a positive value (at most 65536) forces that height */
void set_tile_rows(IPMHandle* self, int val);
int get_tile_rows(IPMHandle* self);
int get_last_tile_rows(IPMHandle* self);                  /* the height the last shoot used */
double get_t_m2l(IPMHandle* self);          /* seconds in the M2L kernels of the last build (MFMA modes) */
long long get_m2l_pairs(IPMHandle* self);   /* M2L (target, source) pairs of the last build (MFMA modes) */
int get_m2l_fp32(IPMHandle* self);          /* 1 if the last build's M2L ran on FP32 MFMA */
void set_kernel(IPMHandle* self, int val);            /* 0: fast shared-corner kernel; 1: reference-faithful FP64 kernel */
int get_kernel(IPMHandle* self);
int get_tree_levels(IPMHandle* self);
ipm_dtype get_alpha_error(IPMHandle* self);
int get_num_ray_threads_1(IPMHandle* self);           /* cells along x1 */
int get_num_ray_threads_2(IPMHandle* self);           /* cells along x2 */
ipm_dtype get_root_half_length(IPMHandle* self);
double get_mean_near_stars(IPMHandle* self);          /* lattice-weighted mean near-field star count */
int get_taylor_smooth(IPMHandle* self);
ipm_dtype get_center_x1(IPMHandle* self);                  /* derived image-plane geometry (ipm.cuh:460-470) */
ipm_dtype get_center_x2(IPMHandle* self);
ipm_dtype get_half_length_x1(IPMHandle* self);
ipm_dtype get_half_length_x2(IPMHandle* self);
ipm_dtype get_ray_half_sep_1(IPMHandle* self);
ipm_dtype get_ray_half_sep_2(IPMHandle* self);
ipm_dtype get_kappa_star_actual(IPMHandle* self);
ipm_dtype get_mean_mass(IPMHandle* self);
ipm_dtype get_mean_mass2(IPMHandle* self);
ipm_dtype get_mean_mass_actual(IPMHandle* self);
ipm_dtype get_mean_mass2_actual(IPMHandle* self);            /* star moments of the realised field (star.cuh:124-161) */
ipm_dtype get_mean_mass2_ln_mass(IPMHandle* self);
ipm_dtype get_mean_mass2_ln_mass_actual(IPMHandle* self);
ipm_dtype get_m_lower_actual(IPMHandle* self);
ipm_dtype get_m_upper_actual(IPMHandle* self);
ipm_dtype get_mu_ave(IPMHandle* self);
/* staged run for benchmarking: run() == prepare + build + shoot(all rows) + finish */
bool ipm_prepare(IPMHandle* self, int verbose);       /* validate, derive, generate stars */
bool ipm_build(IPMHandle* self, int verbose);         /* quadtree + FMM */
bool ipm_shoot(IPMHandle* self, int verbose, long long row0, long long row1);   /* cell rows [row0, row1); row1 < 0: all */
bool ipm_finish(IPMHandle* self, int verbose);        /* histograms + host copies */
long long get_cells_shot(IPMHandle* self);
/* lens equation through the built tree at n points z (re, im pairs) into w;
   mu (may be NULL) receives the magnification, FP64 */
bool ipm_deflect(IPMHandle* self, const double* z, int n, double* w, double* mu);
/* The same as ipm_deflect for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc); it has no counterpart in the reference. It calls no allocating
   routine: the points, deflections and magnifications are staged in a fixed
   array of at most 65536 points (2.6 MB) that each GPU running the library
   holds until the process ends. n above 65536 is refused (false, message on
   stderr). Calls that overlap on one GPU share that array and race. */
bool ipm_deflect_noalloc(IPMHandle* self, const double* z, int n, double* w, double* mu);
/* for checking a window of a large map against a host FP64 re-shoot of just the
   cells that reach it: the cells whose mixed-precision corners' pixel box meets
   columns [x0, x0 + width) and rows [y0, y0 + height) (row 0 at the top), widened
   by margin pixels. Writes at most capacity records of three values (lattice row,
   first column of a 63-cell tile, bit mask of its cells) and returns the number
   found (more than capacity: call again with more room), or -1 before build() */
long long ipm_window_cells(IPMHandle* self, int x0, int y0, int width, int height, double margin, long long* records,
                           long long capacity);
/* the tree as the oracle exports it: geom[3n], info[9n], lists[40n] */
int get_num_nodes(IPMHandle* self);
bool ipm_export_tree(IPMHandle* self, double* geom, int* info, int* lists);
bool ipm_export_coeffs(IPMHandle* self, int which, double* out);   /* 0 multipole, 1 local; 2(p+1) doubles per node */
/* validation and derived geometry only (no device work): num_stars, the
   lattice and the image-plane region become readable (ipm.cuh:370-565) */
bool ipm_derive(IPMHandle* self, int verbose);

/* Many IPM maps in one persistent, graph-captured launch (mixed precision,
   no parity maps). Each instance is an object that has passed ipm_prepare
   and ipm_build; ipm_batch_add allocates its map. ipm_batch_shoot zeroes every
   map, shoots every instance and bins each map into a 128-bin histogram of
   log10(magnification) over [-2, 2] (below/above go to the first/last bin);
   ipm_finish on an instance then copies its map to the host. The first
   shoot times candidate tile heights and resident blocks per CU on a sample
   of the tiles and keeps the fastest. Extension: no reference counterpart. */
typedef struct IPMBatch IPMBatch;
IPMBatch* ipm_batch_create(void);
void ipm_batch_free(IPMBatch* self);
int ipm_batch_add(IPMBatch* self, IPMHandle* instance);        /* index, or -1 */
bool ipm_batch_shoot(IPMBatch* self, int verbose);
double get_batch_t_shoot(IPMBatch* self);                      /* seconds of the last graph launch */
long long get_batch_cells(IPMBatch* self);
int get_batch_size(IPMBatch* self);
bool ipm_batch_histograms(IPMBatch* self, unsigned* out);      /* size x 128 counts */
int get_batch_tile_h(IPMBatch* self);
int get_batch_blocks_per_cu(IPMBatch* self);
double get_batch_t_tune(IPMBatch* self);
bool ipm_batch_set_tuning(IPMBatch* self, int tile_h, int blocks_per_cu);   /* fixed tiles, no autotuning */

#ifdef __cplusplus
}
#endif

#endif
