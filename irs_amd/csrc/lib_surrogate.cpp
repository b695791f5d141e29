// lib_surrogate.so: C ABI of the histogram surrogate (include/lib_surrogate.h).
#include "../../include/lib_surrogate.h"

#include "surrogate.hpp"

#include <iostream>

struct SurrogateHandle
{
	mlens::Surrogate s;
	SurrogateHandle(int max_batch, unsigned seed) : s(max_batch, seed) {}
};

#define SURR_GUARD(expr, fail)                                                         \
	try                                                                                \
	{                                                                                  \
		expr;                                                                          \
	}                                                                                  \
	catch (const mlens::HipError& e)                                                   \
	{                                                                                  \
		std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n";   \
		fail;                                                                          \
	}                                                                                  \
	catch (const std::exception& e)                                                    \
	{                                                                                  \
		std::cerr << "Error. " << e.what() << "\n";                                    \
		fail;                                                                          \
	}

extern "C" {

SurrogateHandle* surrogate_create(int max_batch, unsigned seed)
{
	SURR_GUARD(return new SurrogateHandle(max_batch, seed), return nullptr)
}
void surrogate_free(SurrogateHandle* self) { delete self; }
long long surrogate_num_params(SurrogateHandle* self) { return (long long)self->s.num_params(); }
bool surrogate_set_params(SurrogateHandle* self, const float* params) { SURR_GUARD(self->s.set_params(params); return true, return false) }
bool surrogate_get_params(SurrogateHandle* self, float* params) { SURR_GUARD(self->s.get_params(params); return true, return false) }
bool surrogate_load_data(SurrogateHandle* self, const float* x, const float* t, int n)
{
	SURR_GUARD(self->s.load_data(x, t, n); return true, return false)
}
double surrogate_train_step(SurrogateHandle* self, const int* rows, int m, float lr, float wd)
{
	SURR_GUARD(return self->s.train_step(rows, m, lr, wd), return -1.0)
}
double surrogate_gradients(SurrogateHandle* self, const float* x, const float* t, int n, float* grads)
{
	SURR_GUARD(return self->s.gradients(x, t, n, grads), return -1.0)
}
bool surrogate_predict(SurrogateHandle* self, const float* x, int n, float* probs)
{
	SURR_GUARD(self->s.predict(x, n, probs); return true, return false)
}
double surrogate_time_predict(SurrogateHandle* self, const float* x, int n, int reps)
{
	SURR_GUARD(return self->s.time_predict(x, n, reps), return -1.0)
}

}  // extern "C"
