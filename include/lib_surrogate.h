/*
 * lib_surrogate.h -- C ABI of lib_surrogate.so: a multilayer perceptron that
 * maps lens parameters to the 128-bin histogram of log10(magnification) over
 * [-2, 2] of an IPM map (the histogram ipm_batch_histograms produces), trained
 * and evaluated on the GPU in BF16 with FP32 accumulation.
 *
 * Extension: the reference has no counterpart. Inputs are 6 normalised
 * features per sample (kappa_tot, shear, smooth fraction, mass-function
 * slope, m_lower, m_upper; the caller normalises them); targets and outputs
 * are 128 probabilities per sample. Parameters are laid out layer by layer
 * (6 -> 256 -> 256 -> 256 -> 128, inputs padded to 16): W [out x in] row-major,
 * then b [out]. Calls return false / a negative value with a message on
 * stderr on error.
 */
#ifndef MLENS_LIB_SURROGATE_H
#define MLENS_LIB_SURROGATE_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SurrogateHandle SurrogateHandle;

SurrogateHandle* surrogate_create(int max_batch, unsigned seed);   /* He-uniform weights from Philox(seed) */
void surrogate_free(SurrogateHandle* self);
long long surrogate_num_params(SurrogateHandle* self);
bool surrogate_set_params(SurrogateHandle* self, const float* params);   /* also resets the Adam state */
bool surrogate_get_params(SurrogateHandle* self, float* params);
bool surrogate_load_data(SurrogateHandle* self, const float* x, const float* t, int n);   /* device-resident training set */
double surrogate_train_step(SurrogateHandle* self, const int* rows, int m, float lr, float wd);   /* mean loss; m % 16 == 0 */
double surrogate_gradients(SurrogateHandle* self, const float* x, const float* t, int n, float* grads);   /* loss; n % 16 == 0 */
bool surrogate_predict(SurrogateHandle* self, const float* x, int n, float* probs);
double surrogate_time_predict(SurrogateHandle* self, const float* x, int n, int reps);   /* seconds per histogram */

#ifdef __cplusplus
}
#endif

#endif
