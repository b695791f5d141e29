// The histogram surrogate: a multilayer perceptron mapping lens parameters
// to the 128-bin log10 magnification histogram of a map, trained and run on
// the GPU by the kernels below.
//
// Network: 6 inputs (zero-padded to 16) -> 256 -> 256 -> 256 -> 128 logits,
// SiLU activations, softmax output, cross-entropy against the normalised
// histogram. Products run on the matrix cores in BF16 with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16): every GEMM operand is stored in BF16, every
// result, pre-activation, gradient and the Adam state in FP32; the master
// weights are FP32 and their BF16 copies are refreshed after each step.
#include "surrogate.hpp"

#include "philox.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace mlens
{

namespace
{

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__host__ __device__ inline unsigned short to_bf16(float x)
{
	unsigned u;
	std::memcpy(&u, &x, 4);
	if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
	u += 0x7fffu + ((u >> 16) & 1u);
	return (unsigned short)(u >> 16);
}

/*
C[m][n] (+)= sum_k A(m, k) B(k, n) with A(m, k) = A[m sam + k sak] and
B(k, n) = B[k sbk + n sbn] in BF16, C row-major FP32. M, N, K are multiples of
16. One wave per 16 x 16 output tile; lane l feeds A[l % 16][4 (l / 16) + j]
and B[4 (l / 16) + j][l % 16], j = 0..3, and holds C[4 (l / 16) + r][l % 16].
*/
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const unsigned short* __restrict__ A, long long sam,
	long long sak, const unsigned short* __restrict__ B, long long sbk, long long sbn, float* __restrict__ C, int ldc,
	int accumulate)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tiles_n = N / 16;
	const int tile = blockIdx.x * 4 + wave;
	if (tile >= (M / 16) * tiles_n) return;
	const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
	const int i = lane & 15, kq = lane >> 4;
	f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
	const long long arow = (long long)(tm * 16 + i) * sam, bcol = (long long)(tn * 16 + i) * sbn;
	for (int k0 = 0; k0 < K; k0 += 16)
	{
		bf16x4 a, b;
#pragma unroll
		for (int j = 0; j < 4; j++)
		{
			const long long k = k0 + kq * 4 + j;
			a[j] = (short)A[arow + k * sak];
			b[j] = (short)B[k * sbk + bcol];
		}
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
	}
#pragma unroll
	for (int r = 0; r < 4; r++)
	{
		const long long idx = (long long)(tm * 16 + kq * 4 + r) * ldc + tn * 16 + i;
		C[idx] = accumulate ? C[idx] + acc[r] : acc[r];
	}
}

__device__ inline float sigmoidf(float z) { return 1.0f / (1.0f + expf(-z)); }

// Z = C + b; H = SiLU(Z) in BF16 (hidden layers) or the logits Z (act = 0)
__global__ void k_bias_act(const float* __restrict__ C, const float* __restrict__ b, int M, int N, float* __restrict__ Z,
	unsigned short* __restrict__ H, int act)
{
	const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= (long long)M * N) return;
	const float z = C[g] + b[g % N];
	Z[g] = z;
	if (act) H[g] = to_bf16(z * sigmoidf(z));
}

// one wave per row of 128 logits: P = softmax(Z), loss_m = -sum T log P,
// G = (P - T) / M in FP32 and BF16 (T rows sum to one)
__global__ void k_softmax_xent(const float* __restrict__ Z, const float* __restrict__ T, int M, float* __restrict__ P,
	float* __restrict__ loss, float* __restrict__ G32, unsigned short* __restrict__ G16, float inv_m)
{
	const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (row >= M) return;
	const float* z = Z + (long long)row * SURR_NOUT;
	float z0 = z[lane], z1 = z[lane + 64];
	float mx = fmaxf(z0, z1);
	for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
	float e0 = expf(z0 - mx), e1 = expf(z1 - mx);
	float s = e0 + e1;
	for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
	const float ls = logf(s);
	const float p0 = e0 / s, p1 = e1 / s;
	const float t0 = T ? T[(long long)row * SURR_NOUT + lane] : 0.0f, t1 = T ? T[(long long)row * SURR_NOUT + lane + 64] : 0.0f;
	P[(long long)row * SURR_NOUT + lane] = p0;
	P[(long long)row * SURR_NOUT + lane + 64] = p1;
	if (!T) return;
	// -sum t log p with log p = z - mx - log s
	float l = -(t0 * (z0 - mx - ls) + t1 * (z1 - mx - ls));
	for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
	if (lane == 0) loss[row] = l;
	const float g0 = (p0 - t0) * inv_m, g1 = (p1 - t1) * inv_m;
	G32[(long long)row * SURR_NOUT + lane] = g0;
	G32[(long long)row * SURR_NOUT + lane + 64] = g1;
	G16[(long long)row * SURR_NOUT + lane] = to_bf16(g0);
	G16[(long long)row * SURR_NOUT + lane + 64] = to_bf16(g1);
}

// dZ = dH * SiLU'(Z), SiLU'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_act_bwd(const float* __restrict__ dH, const float* __restrict__ Z, long long n, float* __restrict__ dZ32,
	unsigned short* __restrict__ dZ16)
{
	const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n) return;
	const float z = Z[g], s = sigmoidf(z);
	const float d = dH[g] * (s * (1.0f + z * (1.0f - s)));
	dZ32[g] = d;
	dZ16[g] = to_bf16(d);
}

// db[n] = sum_m dZ[m][n]
__global__ void k_col_sum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= N) return;
	float s = 0.0f;
	for (int m = 0; m < M; m++) s += dZ[(long long)m * N + n];
	db[n] = s;
}

// Adam with decoupled weight decay (Loshchilov & Hutter 2019) on the FP32
// master weights; refreshes the BF16 copy of weight matrices (w16 may be null)
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
	unsigned short* __restrict__ w16, long long n, float lr, float b1, float b2, float eps, float c1, float c2, float wd)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.0f - b1) * gi;
	const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	float wi = w[i] * (1.0f - lr * wd);
	wi -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
	w[i] = wi;
	if (w16) w16[i] = to_bf16(wi);
}

__global__ void k_to_bf16(const float* __restrict__ x, long long n, unsigned short* __restrict__ y)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) y[i] = to_bf16(x[i]);
}

// rows idx[0..M) of the resident training set into the batch buffers
__global__ void k_gather(const int* __restrict__ idx, int M, const unsigned short* __restrict__ X, const float* __restrict__ T,
	unsigned short* __restrict__ Xb, float* __restrict__ Tb)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= M * SURR_NOUT) return;
	const int r = g / SURR_NOUT, c = g - r * SURR_NOUT;
	const long long src = idx[r];
	Tb[g] = T[src * SURR_NOUT + c];
	if (c < SURR_PIN) Xb[(long long)r * SURR_PIN + c] = X[src * SURR_PIN + c];
}

inline unsigned blocks(long long n, int t) { return unsigned((n + t - 1) / t); }

}  // namespace

const int Surrogate::dims_[SURR_LAYERS + 1] = {SURR_PIN, SURR_HIDDEN, SURR_HIDDEN, SURR_HIDDEN, SURR_NOUT};

Surrogate::Surrogate(int max_batch, uint32_t seed) : max_batch_((max_batch + 15) / 16 * 16)
{
	MLENS_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
	// parameter layout: W_l [out x in] then b_l [out] for l = 0..3
	size_t off = 0;
	for (int l = 0; l < SURR_LAYERS; l++)
	{
		w_off_[l] = off;
		off += size_t(dims_[l + 1]) * dims_[l];
		b_off_[l] = off;
		off += dims_[l + 1];
	}
	nparams_ = off;
	w_.alloc(nparams_);
	g_.alloc(nparams_);
	m_.alloc(nparams_);
	v_.alloc(nparams_);
	w16_.alloc(nparams_);
	// He-uniform weights from Philox (key = seed), zero biases; padded input columns stay zero
	std::vector<float> w(nparams_, 0.0f);
	for (int l = 0; l < SURR_LAYERS; l++)
	{
		const int fin = l == 0 ? SURR_NIN : dims_[l];
		const float a = std::sqrt(6.0f / float(fin));
		for (int o = 0; o < dims_[l + 1]; o++)
			for (int i = 0; i < fin; i++)
			{
				const uint64_t k = w_off_[l] + size_t(o) * dims_[l] + i;
				uint32_t ctr[4] = {uint32_t(k), uint32_t(k >> 32), 7u, 0u}, key[2] = {seed, 0x5eed5eedu}, r[4];
				philox4x32_10(ctr, key, r);
				const float u = float(r[0] >> 8) * (1.0f / 16777216.0f);
				w[w_off_[l] + size_t(o) * dims_[l] + i] = a * (2.0f * u - 1.0f);
			}
	}
	set_params(w.data());
	const size_t mb = size_t(max_batch_);
	xb_.alloc(mb * SURR_PIN);
	tb_.alloc(mb * SURR_NOUT);
	idx_.alloc(mb);
	for (int l = 0; l < SURR_LAYERS; l++)
	{
		z_[l].alloc(mb * dims_[l + 1]);
		h_[l].alloc(mb * dims_[l + 1]);
	}
	c_.alloc(mb * SURR_HIDDEN);
	p_.alloc(mb * SURR_NOUT);
	loss_.alloc(mb);
	d32_.alloc(mb * SURR_HIDDEN);
	d16_.alloc(mb * SURR_HIDDEN);
	dh_.alloc(mb * SURR_HIDDEN);
}

Surrogate::~Surrogate()
{
	if (stream_)
	{
		(void)hipStreamSynchronize(stream_);
		(void)hipStreamDestroy(stream_);
	}
}

void Surrogate::set_params(const float* host)
{
	MLENS_HIP_CHECK(hipMemcpy(w_.ptr, host, nparams_ * sizeof(float), hipMemcpyHostToDevice));
	k_to_bf16<<<blocks(nparams_, 256), 256, 0, stream_>>>(w_.ptr, (long long)nparams_, w16_.ptr);
	MLENS_HIP_CHECK(hipMemsetAsync(m_.ptr, 0, nparams_ * sizeof(float), stream_));
	MLENS_HIP_CHECK(hipMemsetAsync(v_.ptr, 0, nparams_ * sizeof(float), stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	step_ = 0;
}

void Surrogate::get_params(float* host) const
{
	MLENS_HIP_CHECK(hipMemcpy(host, w_.ptr, nparams_ * sizeof(float), hipMemcpyDeviceToHost));
}

void Surrogate::gemm(int M, int N, int K, const unsigned short* A, long long sam, long long sak, const unsigned short* B,
	long long sbk, long long sbn, float* C, int ldc, bool accumulate)
{
	if (M % 16 || N % 16 || K % 16) throw HipError("surrogate GEMM sizes must be multiples of 16", __FILE__, __LINE__);
	const long long tiles = (long long)(M / 16) * (N / 16);
	k_gemm<<<blocks(tiles, 4), 256, 0, stream_>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, accumulate ? 1 : 0);
	MLENS_HIP_CHECK(hipGetLastError());
}

// forward pass of the batch in xb_ (M rows, M a multiple of 16); logits in z_[3]
void Surrogate::forward(int M)
{
	const unsigned short* in = xb_.ptr;
	for (int l = 0; l < SURR_LAYERS; l++)
	{
		const int K = dims_[l], N = dims_[l + 1];
		// Z = H_prev W^T: A(m, k) = in[m K + k], B(k, n) = W[n K + k]
		gemm(M, N, K, in, K, 1, w16_.ptr + w_off_[l], 1, K, c_.ptr, N, false);
		k_bias_act<<<blocks((long long)M * N, 256), 256, 0, stream_>>>(c_.ptr, w_.ptr + b_off_[l], M, N, z_[l].ptr, h_[l].ptr,
			l + 1 < SURR_LAYERS ? 1 : 0);
		in = h_[l].ptr;
	}
}

// backward pass after forward(M) with the targets in tb_: parameter gradients in g_; returns the mean loss
double Surrogate::backward(int M)
{
	k_softmax_xent<<<blocks(M, 4), 256, 0, stream_>>>(z_[3].ptr, tb_.ptr, M, p_.ptr, loss_.ptr, d32_.ptr, d16_.ptr, 1.0f / float(M));
	for (int l = SURR_LAYERS - 1; l >= 0; l--)
	{
		const int K = dims_[l], N = dims_[l + 1];
		const unsigned short* in = l == 0 ? xb_.ptr : h_[l - 1].ptr;
		// dZ_l is in d32_ / d16_ (M x N)
		// dW_l = dZ^T H_prev: C(n, k) = sum_m dZ[m N + n] in[m K + k]
		gemm(N, K, M, d16_.ptr, 1, N, in, K, 1, g_.ptr + w_off_[l], K, false);
		k_col_sum<<<blocks(N, 256), 256, 0, stream_>>>(d32_.ptr, M, N, g_.ptr + b_off_[l]);
		if (l == 0) break;
		// dH_prev = dZ W: C(m, k) = sum_n dZ[m N + n] W[n K + k]
		gemm(M, K, N, d16_.ptr, N, 1, w16_.ptr + w_off_[l], K, 1, dh_.ptr, K, false);
		k_act_bwd<<<blocks((long long)M * K, 256), 256, 0, stream_>>>(dh_.ptr, z_[l - 1].ptr, (long long)M * K, d32_.ptr, d16_.ptr);
	}
	MLENS_HIP_CHECK(hipGetLastError());
	std::vector<float> loss(M);
	MLENS_HIP_CHECK(hipMemcpyAsync(loss.data(), loss_.ptr, M * sizeof(float), hipMemcpyDeviceToHost, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	double s = 0;
	for (float x : loss) s += x;
	return s / M;
}

void Surrogate::load_data(const float* x, const float* t, int n)
{
	// inputs padded to SURR_PIN columns and stored in BF16; targets FP32
	std::vector<float> xp(size_t(n) * SURR_PIN, 0.0f);
	for (int r = 0; r < n; r++)
		for (int c = 0; c < SURR_NIN; c++) xp[size_t(r) * SURR_PIN + c] = x[size_t(r) * SURR_NIN + c];
	DeviceBuffer<float> tmp;
	tmp.alloc(xp.size());
	MLENS_HIP_CHECK(hipMemcpy(tmp.ptr, xp.data(), xp.size() * sizeof(float), hipMemcpyHostToDevice));
	xall_.alloc(xp.size());
	k_to_bf16<<<blocks((long long)xp.size(), 256), 256, 0, stream_>>>(tmp.ptr, (long long)xp.size(), xall_.ptr);
	tall_.alloc(size_t(n) * SURR_NOUT);
	MLENS_HIP_CHECK(hipMemcpyAsync(tall_.ptr, t, size_t(n) * SURR_NOUT * sizeof(float), hipMemcpyHostToDevice, stream_));
	MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	ndata_ = n;
}

double Surrogate::train_step(const int* rows, int M, float lr, float wd)
{
	if (M <= 0 || M > max_batch_ || M % 16) throw HipError("surrogate batch must be a positive multiple of 16 within the maximum", __FILE__, __LINE__);
	for (int r = 0; r < M; r++)
		if (rows[r] < 0 || rows[r] >= ndata_) throw HipError("surrogate row index outside the training set", __FILE__, __LINE__);
	MLENS_HIP_CHECK(hipMemcpyAsync(idx_.ptr, rows, M * sizeof(int), hipMemcpyHostToDevice, stream_));
	k_gather<<<blocks((long long)M * SURR_NOUT, 256), 256, 0, stream_>>>(idx_.ptr, M, xall_.ptr, tall_.ptr, xb_.ptr, tb_.ptr);
	forward(M);
	const double loss = backward(M);
	step_++;
	const float b1 = 0.9f, b2 = 0.999f;
	const float c1 = 1.0f / (1.0f - std::pow(b1, float(step_))), c2 = 1.0f / (1.0f - std::pow(b2, float(step_)));
	for (int l = 0; l < SURR_LAYERS; l++)
	{
		const long long nw = (long long)dims_[l + 1] * dims_[l], nb = dims_[l + 1];
		k_adam<<<blocks(nw, 256), 256, 0, stream_>>>(w_.ptr + w_off_[l], g_.ptr + w_off_[l], m_.ptr + w_off_[l],
			v_.ptr + w_off_[l], w16_.ptr + w_off_[l], nw, lr, b1, b2, 1e-8f, c1, c2, wd);
		k_adam<<<blocks(nb, 256), 256, 0, stream_>>>(w_.ptr + b_off_[l], g_.ptr + b_off_[l], m_.ptr + b_off_[l],
			v_.ptr + b_off_[l], nullptr, nb, lr, b1, b2, 1e-8f, c1, c2, 0.0f);
	}
	MLENS_HIP_CHECK(hipGetLastError());
	// the zero-padded input columns of W_0 get zero gradients and stay zero
	return loss;
}

double Surrogate::gradients(const float* x, const float* t, int n, float* grads)
{
	std::vector<int> rows(n);
	for (int i = 0; i < n; i++) rows[i] = i;
	load_data(x, t, n);
	const int M = (n + 15) / 16 * 16;
	if (M != n) throw HipError("gradient batch must be a multiple of 16", __FILE__, __LINE__);
	MLENS_HIP_CHECK(hipMemcpyAsync(idx_.ptr, rows.data(), n * sizeof(int), hipMemcpyHostToDevice, stream_));
	k_gather<<<blocks((long long)M * SURR_NOUT, 256), 256, 0, stream_>>>(idx_.ptr, M, xall_.ptr, tall_.ptr, xb_.ptr, tb_.ptr);
	forward(M);
	const double loss = backward(M);
	MLENS_HIP_CHECK(hipMemcpy(grads, g_.ptr, nparams_ * sizeof(float), hipMemcpyDeviceToHost));
	return loss;
}

void Surrogate::predict(const float* x, int n, float* probs)
{
	for (int r0 = 0; r0 < n; r0 += max_batch_)
	{
		const int rows = std::min(max_batch_, n - r0), M = (rows + 15) / 16 * 16;
		std::vector<unsigned short> xp(size_t(M) * SURR_PIN, 0);
		for (int r = 0; r < rows; r++)
			for (int c = 0; c < SURR_NIN; c++) xp[size_t(r) * SURR_PIN + c] = to_bf16(x[size_t(r0 + r) * SURR_NIN + c]);
		MLENS_HIP_CHECK(hipMemcpyAsync(xb_.ptr, xp.data(), xp.size() * sizeof(unsigned short), hipMemcpyHostToDevice, stream_));
		forward(M);
		k_softmax_xent<<<blocks(M, 4), 256, 0, stream_>>>(z_[3].ptr, nullptr, M, p_.ptr, nullptr, nullptr, nullptr, 0.0f);
		MLENS_HIP_CHECK(hipGetLastError());
		MLENS_HIP_CHECK(hipMemcpyAsync(probs + size_t(r0) * SURR_NOUT, p_.ptr, size_t(rows) * SURR_NOUT * sizeof(float),
			hipMemcpyDeviceToHost, stream_));
		MLENS_HIP_CHECK(hipStreamSynchronize(stream_));
	}
}

double Surrogate::time_predict(const float* x, int n, int reps)
{
	std::vector<float> probs(size_t(n) * SURR_NOUT);
	predict(x, n, probs.data());
	hipEvent_t e0, e1;
	MLENS_HIP_CHECK(hipEventCreate(&e0));
	MLENS_HIP_CHECK(hipEventCreate(&e1));
	const int M = (std::min(n, max_batch_) + 15) / 16 * 16;
	MLENS_HIP_CHECK(hipEventRecord(e0, stream_));
	for (int r = 0; r < reps; r++)
	{
		forward(M);
		k_softmax_xent<<<blocks(M, 4), 256, 0, stream_>>>(z_[3].ptr, nullptr, M, p_.ptr, nullptr, nullptr, nullptr, 0.0f);
	}
	MLENS_HIP_CHECK(hipEventRecord(e1, stream_));
	MLENS_HIP_CHECK(hipEventSynchronize(e1));
	float ms = 0;
	MLENS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return ms / 1000.0 / reps / M;   // seconds per predicted histogram
}

}  // namespace mlens
