/*
 * lib_ncc.h -- C ABI of lib_ncc.so: the number of caustic crossings over a
 * source-plane pixel grid. Values are double, as the reference builds its
 * lib_ncc.so (compile:11).
 *
 * Drop-in replacement for the reference's src/lib_ncc.cu:12-46: the same 26
 * unprefixed extern "C" symbols with the same argument meaning, so the
 * reference's own ctypes binding (microlensing/NCC/lib_ncc.py:5-62) loads it
 * unchanged.
 *
 * run() reads <infile_prefix>ccf_caustics.bin (int32 rows, int32 columns,
 * then the complex caustic points in double or float), validates its
 * parameters before it touches the GPU and returns false with a message on
 * stderr on any error, including a map with a negative count (ncc.cuh:321-326).
 * The object returned by init() is opaque; the array returned by
 * get_num_crossings() is host memory owned by the object, valid until that
 * object's next run() or ncc_free().
 *
 * Map layout (ncc.cuh:425-433): num_pixels_y2 rows of num_pixels_y1 int32
 * counts, row 0 at the top (largest y2). Each count is the maximum, over the
 * 2^over_sample x 2^over_sample sub-pixels of the pixel, of the number of
 * caustics enclosing the sub-pixel's centre.
 *
 * Entries marked "extension" have no counterpart in the reference.
 */
#ifndef MLENS_LIB_NCC_H
#define MLENS_LIB_NCC_H

#include <stdbool.h>

typedef double ncc_dtype;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NCCHandle NCCHandle;

/* lib_ncc.cu:15 */
NCCHandle* init(void);

/* setters, lib_ncc.cu:17-27 */
void set_infile_prefix(NCCHandle* self, const char* val);
void set_center_y1(NCCHandle* self, ncc_dtype val);
void set_center_y2(NCCHandle* self, ncc_dtype val);
void set_half_length_y1(NCCHandle* self, ncc_dtype val);
void set_half_length_y2(NCCHandle* self, ncc_dtype val);
void set_num_pixels_y1(NCCHandle* self, int val);
void set_num_pixels_y2(NCCHandle* self, int val);
void set_over_sample(NCCHandle* self, int val);
void set_write_maps(NCCHandle* self, int val);
void set_write_histograms(NCCHandle* self, int val);
void set_outfile_prefix(NCCHandle* self, const char* val);

/* getters, lib_ncc.cu:29-39 */
const char* get_infile_prefix(NCCHandle* self);
ncc_dtype get_center_y1(NCCHandle* self);
ncc_dtype get_center_y2(NCCHandle* self);
ncc_dtype get_half_length_y1(NCCHandle* self);
ncc_dtype get_half_length_y2(NCCHandle* self);
int get_num_pixels_y1(NCCHandle* self);
int get_num_pixels_y2(NCCHandle* self);
int get_over_sample(NCCHandle* self);
int get_write_maps(NCCHandle* self);
int get_write_histograms(NCCHandle* self);
const char* get_outfile_prefix(NCCHandle* self);

/* lib_ncc.cu:41 */
int* get_num_crossings(NCCHandle* self);   /* num_pixels_y2 x num_pixels_y1, null before run() */

/* lib_ncc.cu:43-44 */
bool run(NCCHandle* self, int verbose);
bool save(NCCHandle* self, int verbose);

/* extensions */
void ncc_free(NCCHandle* self);            /* the reference leaks its objects */
double get_t_ncc(NCCHandle* self);         /* seconds counting the crossings */
double get_t_reduce(NCCHandle* self);      /* seconds reducing the over-sampled grid */

#ifdef __cplusplus
}
#endif

#endif
