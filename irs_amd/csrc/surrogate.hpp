// The histogram surrogate network (see surrogate.hip).
#pragma once

#include "engine.hpp"

namespace mlens
{

constexpr int SURR_NIN = 6;        // kappa_tot, shear, smooth fraction, mass-function slope, m_lower, m_upper
constexpr int SURR_PIN = 16;       // inputs padded to one MFMA K step
constexpr int SURR_HIDDEN = 256;
constexpr int SURR_NOUT = 128;     // log10 magnification bins over [-2, 2]
constexpr int SURR_LAYERS = 4;

class Surrogate
{
public:
	Surrogate(int max_batch, uint32_t seed);
	~Surrogate();
	Surrogate(const Surrogate&) = delete;
	Surrogate& operator=(const Surrogate&) = delete;

	size_t num_params() const { return nparams_; }
	void set_params(const float* host);   // FP32 master weights (layer by layer: W [out x in], then b); resets Adam
	void get_params(float* host) const;

	void load_data(const float* x, const float* t, int n);            // x: n x 6 (normalised), t: n x 128 (rows sum to 1)
	double train_step(const int* rows, int m, float lr, float wd);   // one Adam step on rows of the loaded set; mean loss
	double gradients(const float* x, const float* t, int n, float* grads);   // loss and dloss/dparams of one batch (n % 16 == 0)
	void predict(const float* x, int n, float* probs);                // n x 128 softmax outputs
	double time_predict(const float* x, int n, int reps);             // seconds per predicted histogram on the device

private:
	void gemm(int M, int N, int K, const unsigned short* A, long long sam, long long sak, const unsigned short* B,
		long long sbk, long long sbn, float* C, int ldc, bool accumulate);
	void forward(int M);
	double backward(int M);

	static const int dims_[SURR_LAYERS + 1];
	int max_batch_;
	hipStream_t stream_ = nullptr;
	size_t nparams_ = 0;
	size_t w_off_[SURR_LAYERS], b_off_[SURR_LAYERS];
	int step_ = 0;
	int ndata_ = 0;
	DeviceBuffer<float> w_, g_, m_, v_;
	DeviceBuffer<unsigned short> w16_;
	DeviceBuffer<unsigned short> xall_, xb_;
	DeviceBuffer<float> tall_, tb_;
	DeviceBuffer<int> idx_;
	DeviceBuffer<float> z_[SURR_LAYERS];
	DeviceBuffer<unsigned short> h_[SURR_LAYERS];
	DeviceBuffer<float> c_, p_, loss_, d32_, dh_;
	DeviceBuffer<unsigned short> d16_;
};

}  // namespace mlens
