/*
 * lib_ccf.h -- C ABI of lib_ccf.so: critical curves and caustics of a star
 * field. Values are double, as the reference builds its lib_ccf.so (compile:8).
 *
 * Drop-in replacement for the reference's src/lib_ccf.cu:14-70: the same 53
 * unprefixed extern "C" symbols with the same argument meaning, so the
 * reference's own ctypes binding (microlensing/CCF/lib_ccf.py:5-116) loads it
 * unchanged.
 *
 * The object returned by init() is opaque. Arrays returned by the getters are
 * host memory owned by the object, valid until that object's next run() or
 * ccf_free(). run() validates its parameters before it touches the GPU and
 * returns false with a message on stderr on any error. The roots are refined
 * in FP64.
 *
 * Array layouts (lib_ccf.cu:62-65, ccf.cuh:1180-1230): critical curves and
 * caustics are (num_roots * num_branches) rows of (num_phi / num_branches + 1)
 * complex points (x1, x2 interleaved); row r * num_branches + b is root r on
 * branch b, its columns the branch's phases in increasing order. Magnification
 * length scales have the same shape with one real value per point.
 *
 * Entries marked "extension" have no counterpart in the reference.
 */
#ifndef MLENS_LIB_CCF_H
#define MLENS_LIB_CCF_H

#include <stdbool.h>

typedef double ccf_dtype;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct CCFHandle CCFHandle;

/* lib_ccf.cu:16 */
CCFHandle* init(void);

/* setters, lib_ccf.cu:18-38 */
void set_kappa_tot(CCFHandle* self, ccf_dtype val);
void set_shear(CCFHandle* self, ccf_dtype val);
void set_kappa_star(CCFHandle* self, ccf_dtype val);
void set_theta_star(CCFHandle* self, ccf_dtype val);
void set_mass_function(CCFHandle* self, const char* val);
void set_m_solar(CCFHandle* self, ccf_dtype val);
void set_m_lower(CCFHandle* self, ccf_dtype val);
void set_m_upper(CCFHandle* self, ccf_dtype val);
void set_rectangular(CCFHandle* self, int val);
void set_approx(CCFHandle* self, int val);
void set_safety_scale(CCFHandle* self, ccf_dtype val);
void set_num_stars(CCFHandle* self, int val);
void set_starfile(CCFHandle* self, const char* val);
void set_num_phi(CCFHandle* self, int val);
void set_num_branches(CCFHandle* self, int val);
void set_random_seed(CCFHandle* self, int val);
void set_write_stars(CCFHandle* self, int val);
void set_write_critical_curves(CCFHandle* self, int val);
void set_write_caustics(CCFHandle* self, int val);
void set_write_mu_length_scales(CCFHandle* self, int val);
void set_outfile_prefix(CCFHandle* self, const char* val);

/* getters, lib_ccf.cu:40-60 */
ccf_dtype get_kappa_tot(CCFHandle* self);
ccf_dtype get_shear(CCFHandle* self);
ccf_dtype get_kappa_star(CCFHandle* self);
ccf_dtype get_theta_star(CCFHandle* self);
const char* get_mass_function(CCFHandle* self);
ccf_dtype get_m_solar(CCFHandle* self);
ccf_dtype get_m_lower(CCFHandle* self);
ccf_dtype get_m_upper(CCFHandle* self);
int get_rectangular(CCFHandle* self);
int get_approx(CCFHandle* self);
ccf_dtype get_safety_scale(CCFHandle* self);
int get_num_stars(CCFHandle* self);
const char* get_starfile(CCFHandle* self);
int get_num_phi(CCFHandle* self);
int get_num_branches(CCFHandle* self);
int get_random_seed(CCFHandle* self);
int get_write_stars(CCFHandle* self);
int get_write_critical_curves(CCFHandle* self);
int get_write_caustics(CCFHandle* self);
int get_write_mu_length_scales(CCFHandle* self);
const char* get_outfile_prefix(CCFHandle* self);

/* derived values and results, lib_ccf.cu:62-69 */
int get_num_roots(CCFHandle* self);
ccf_dtype get_corner_x1(CCFHandle* self);          /* (R, 0) for a circular field */
ccf_dtype get_corner_x2(CCFHandle* self);
ccf_dtype* get_stars(CCFHandle* self);             /* num_stars x (x1, x2, m) */
ccf_dtype* get_critical_curves(CCFHandle* self);
ccf_dtype* get_caustics(CCFHandle* self);          /* null unless write_caustics */
ccf_dtype* get_mu_length_scales(CCFHandle* self);  /* null unless write_mu_length_scales */
double get_t_ccs(CCFHandle* self);                 /* seconds of the phase sweep */

/* lib_ccf.cu:71-72 */
bool run(CCFHandle* self, int verbose);
bool save(CCFHandle* self, int verbose);

/* extensions */
void ccf_free(CCFHandle* self);                    /* the reference leaks its objects */
double get_t_init_roots(CCFHandle* self);          /* seconds refining the starting roots */
double get_max_error(CCFHandle* self);             /* largest error bound on 1/mu */
int get_expansion_order(CCFHandle* self);
int get_tree_levels(CCFHandle* self);

#ifdef __cplusplus
}
#endif

#endif
