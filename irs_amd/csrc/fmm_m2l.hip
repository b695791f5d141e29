// M2L as batched GEMMs on the matrix cores (fmm.cuh:281-374 restated as
// matrix products).
//
// A source f in the M2L list of a target t sits at the target's level, at
// dz = (c_f - c_t) / h = 2 (a + i b) with integers |a|, |b| <= 3 and
// max(|a|, |b|) >= 2: 40 offsets. In node-normalised units the translation
// is linear in the source multipole b and depends only on the offset:
//
//   c_0 += sum_{k>=1} b_k (-dz)^-k + b_0 log(-dz) + b_0 log h
//   c_l += sum_{k>=1} C(l+k-1, k-1) b_k (-dz)^-k dz^-l - b_0 / (l dz^l),  l >= 1
//
// so c += M_o b with 40 fixed (p+1) x (p+1) complex matrices M_o, plus the
// level-dependent b_0 log h term. Complex products run as real products of
// the stacked form [Mr -Mi; Mi Mr] [b_re; b_im], padded to a multiple of 16.
//
// One wave owns 16 targets (the columns of a 16 x 16 output tile); for each
// offset it gathers the 16 source multipoles at that offset (zero where a
// target has none) and accumulates M_o b over all offsets in its registers,
// so no target is written by two waves and no atomics are needed. The block's
// four waves share each offset matrix through LDS. The MFMA shape is
// 16 x 16 x 4 in FP32 (v_mfma_f32_16x16x4_f32) or FP64
// (v_mfma_f64_16x16x4_f64); both take A[i][k] from lane i + 16 k and B[k][j]
// from lane j + 16 k. Lane l gets back D[4 (l / 16) + r][l % 16] in FP32 and
// D[(l / 16) + 4 r][l % 16] in FP64, r = 0..3 (checked on the device).
#include "engine.hpp"
#include "device_math.hpp"

#include <complex>
#include <vector>

namespace mlens
{

namespace
{

constexpr int M2L_OFFSETS = 40;
constexpr int M2L_WAVES = 4;
constexpr int M2L_TARGETS_PER_WAVE = 16;

inline int m2l_kp(int p) { return (2 * (p + 1) + 15) / 16 * 16; }

// offset (a, b), |a|, |b| <= 3, to its index 0..39, or -1 for the 3 x 3 near block
__host__ __device__ inline int m2l_offset_index(int a, int b)
{
	if (a < -3 || a > 3 || b < -3 || b > 3) return -1;
	if (a >= -1 && a <= 1 && b >= -1 && b <= 1) return -1;
	int idx = (b + 3) * 7 + (a + 3);   // 0..48 over the 7 x 7 block
	// drop the 9 near entries that precede idx
	int before = 0;
	for (int bb = -1; bb <= 1; bb++)
		for (int aa = -1; aa <= 1; aa++)
			if ((bb + 3) * 7 + (aa + 3) < idx) before++;
	return idx - before;
}

template <typename T>
struct MfmaVec;
template <>
struct MfmaVec<float>
{
	typedef float v4 __attribute__((ext_vector_type(4)));
	static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
	static __device__ __forceinline__ int row(int kq, int r) { return kq * 4 + r; }
};
template <>
struct MfmaVec<double>
{
	typedef double v4 __attribute__((ext_vector_type(4)));
	static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
	static __device__ __forceinline__ int row(int kq, int r) { return kq + 4 * r; }
};

// src[o * n + (t - l0)] = the M2L source of target t at offset o, else -1
__global__ void k_m2l_table(const double* cx, const double* cy, const double* hl, const int* nm2l, const int* m2l, int l0,
	int n, int* src, unsigned long long* pairs, int* bad)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const int i = l0 + t;
	const double h2 = 2 * hl[i];
	const int cnt = nm2l[i];
	for (int j = 0; j < cnt; j++)
	{
		const int f = m2l[27 * i + j];
		const int a = int(lrint((cx[f] - cx[i]) / h2)), b = int(lrint((cy[f] - cy[i]) / h2));
		const int o = m2l_offset_index(a, b);
		if (o < 0 || f < l0 || f >= l0 + n)
		{
			atomicOr(bad, 1);
			continue;
		}
		src[(long long)o * n + t] = f - l0;
	}
	if (cnt > 0) atomicAdd(pairs, (unsigned long long)cnt);
}

__global__ void k_fill_int(int* a, long long n, int v)
{
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = v;
}

// real-stacked multipoles of the level's nodes: row k = Re b_k, row P + k = Im b_k
template <typename T>
__global__ void k_mstack(const double* mult, int l0, int n, int p, int kp, T* ms)
{
	long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= (long long)n * kp) return;
	const int t = int(g / kp), r = int(g % kp), P = p + 1;
	const double* b = mult + (long long)(l0 + t) * P * 2;
	T v = 0;
	if (r < P)
		v = T(b[2 * r]);
	else if (r < 2 * P)
		v = T(b[2 * (r - P) + 1]);
	ms[g] = v;
}

template <typename T>
__global__ __launch_bounds__(64 * M2L_WAVES) void k_m2l_mfma(const T* __restrict__ At, const int* __restrict__ src,
	const T* __restrict__ ms, const double* __restrict__ mult, double* local, const double* hl, int l0, int n, int p, int kp)
{
	extern __shared__ char smem_raw[];
	T* As = reinterpret_cast<T*>(smem_raw);   // kp x kp, transposed: As[k * kp + row]
	using V = typename MfmaVec<T>::v4;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int t0 = (blockIdx.x * M2L_WAVES + wave) * M2L_TARGETS_PER_WAVE;
	const int col = lane & 15, kq = lane >> 4;
	const int t = t0 + col;
	const int ntile = kp / 16;   // row tiles (<= 4)
	V acc[4];
#pragma unroll
	for (int r = 0; r < 4; r++) acc[r] = V{0, 0, 0, 0};
	for (int o = 0; o < M2L_OFFSETS; o++)
	{
		__syncthreads();
		const T* Ao = At + (long long)o * kp * kp;
		for (int i = threadIdx.x; i < kp * kp; i += blockDim.x) As[i] = Ao[i];
		__syncthreads();
		const int s = (t < n) ? src[(long long)o * n + t] : -1;
		if (__ballot(s >= 0) == 0) continue;
		const T* bs = ms + (long long)(s >= 0 ? s : 0) * kp;
		for (int ks = 0; ks < kp / 4; ks++)
		{
			const int k = ks * 4 + kq;
			const T b = s >= 0 ? bs[k] : T(0);
#pragma unroll
			for (int rt = 0; rt < 4; rt++)
				if (rt < ntile) acc[rt] = MfmaVec<T>::mfma(As[k * kp + rt * 16 + col], b, acc[rt]);
		}
	}
	if (t >= n) return;
	const int P = p + 1;
	// b_0 log h in FP64 from the FP64 multipoles, added with row 0 (real) and
	// row P (imaginary) by the lanes that own those rows
	const double lh = log(hl[l0 + t]);
	double sre = 0, sim = 0;
	for (int o = 0; o < M2L_OFFSETS; o++)
	{
		const int sidx = src[(long long)o * n + t];
		if (sidx >= 0)
		{
			sre += mult[(long long)(l0 + sidx) * P * 2];
			sim += mult[(long long)(l0 + sidx) * P * 2 + 1];
		}
	}
	double* c = local + (long long)(l0 + t) * P * 2;
#pragma unroll
	for (int rt = 0; rt < 4; rt++)
	{
		if (rt >= ntile) break;
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			const int row = rt * 16 + MfmaVec<T>::row(kq, r);
			if (row < P)
				c[2 * row] += double(acc[rt][r]) + (row == 0 ? sre * lh : 0.0);
			else if (row < 2 * P)
				c[2 * (row - P) + 1] += double(acc[rt][r]) + (row == P ? sim * lh : 0.0);
		}
	}
}

}  // namespace

// the 40 offset matrices for order p, real-stacked, transposed and padded
void Engine::m2l_prepare(int p, bool f32)
{
	if (m2l_ready_p_ == p && m2l_ready_f32_ == f32) return;
	const int P = p + 1, kp = m2l_kp(p);
	// exact binomials C(n, k), n <= 2p
	std::vector<std::vector<unsigned long long>> C(2 * P + 1);
	for (int nr = 0; nr <= 2 * P; nr++)
	{
		C[nr].assign(nr + 1, 1ull);
		for (int k = 1; k < nr; k++) C[nr][k] = C[nr - 1][k - 1] + C[nr - 1][k];
	}
	std::vector<double> A(size_t(M2L_OFFSETS) * kp * kp, 0.0);
	for (int b = -3; b <= 3; b++)
		for (int a = -3; a <= 3; a++)
		{
			const int o = m2l_offset_index(a, b);
			if (o < 0) continue;
			const std::complex<double> dz(2.0 * a, 2.0 * b), mdz = -dz, imdz = 1.0 / mdz;
			std::vector<std::complex<double>> ipow(P), dzinv(P);
			ipow[0] = 1.0;
			dzinv[0] = 1.0;
			for (int k = 1; k < P; k++)
			{
				ipow[k] = ipow[k - 1] * imdz;
				dzinv[k] = dzinv[k - 1] / dz;
			}
			double* Ao = A.data() + size_t(o) * kp * kp;
			auto put = [&](int l, int k, std::complex<double> m) {
				// stacked [Mr -Mi; Mi Mr], stored transposed: At[col][row]
				Ao[size_t(k) * kp + l] = m.real();
				Ao[size_t(P + k) * kp + l] = -m.imag();
				Ao[size_t(k) * kp + P + l] = m.imag();
				Ao[size_t(P + k) * kp + P + l] = m.real();
			};
			put(0, 0, std::log(mdz));
			for (int k = 1; k < P; k++) put(0, k, ipow[k]);
			for (int l = 1; l < P; l++)
			{
				put(l, 0, -dzinv[l] / double(l));
				for (int k = 1; k < P; k++) put(l, k, double(C[l + k - 1][k - 1]) * ipow[k] * dzinv[l]);
			}
		}
	if (f32)
	{
		std::vector<float> Af(A.begin(), A.end());
		m2l_af_.alloc(Af.size());
		MLENS_HIP_CHECK(hipMemcpy(m2l_af_.ptr, Af.data(), Af.size() * sizeof(float), hipMemcpyHostToDevice));
	}
	else
	{
		m2l_ad_.alloc(A.size());
		MLENS_HIP_CHECK(hipMemcpy(m2l_ad_.ptr, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice));
	}
	m2l_ready_p_ = p;
	m2l_ready_f32_ = f32;
}

void Engine::m2l_level(int lev, int p, bool f32, hipEvent_t e0, hipEvent_t e1)
{
	const int l0 = level_start_[lev], n = level_start_[lev + 1] - l0;
	if (n <= 0)
	{
		MLENS_HIP_CHECK(hipEventRecord(e0, stream_));
		MLENS_HIP_CHECK(hipEventRecord(e1, stream_));
		return;
	}
	const int kp = m2l_kp(p);
	const long long nsrc = (long long)M2L_OFFSETS * n;
	if ((long long)m2l_src_.count < nsrc) m2l_src_.alloc(size_t(nsrc));
	k_fill_int<<<unsigned((nsrc + 255) / 256), 256, 0, stream_>>>(m2l_src_.ptr, nsrc, -1);
	int* bad = reinterpret_cast<int*>(m2l_count_.ptr + 1);
	k_m2l_table<<<unsigned((n + 255) / 256), 256, 0, stream_>>>(ncx_.ptr, ncy_.ptr, nhl_.ptr, nnm2l_.ptr, nm2l_.ptr, l0, n,
		m2l_src_.ptr, m2l_count_.ptr, bad);
	const long long nms = (long long)n * kp;
	const unsigned blocks = unsigned((n + M2L_WAVES * M2L_TARGETS_PER_WAVE - 1) / (M2L_WAVES * M2L_TARGETS_PER_WAVE));
	if (f32)
	{
		if ((long long)mstack_f_.count < nms) mstack_f_.alloc(size_t(nms));
		k_mstack<float><<<unsigned((nms + 255) / 256), 256, 0, stream_>>>(mult_.ptr, l0, n, p, kp, mstack_f_.ptr);
		MLENS_HIP_CHECK(hipEventRecord(e0, stream_));
		k_m2l_mfma<float><<<blocks, 64 * M2L_WAVES, size_t(kp) * kp * sizeof(float), stream_>>>(m2l_af_.ptr, m2l_src_.ptr,
			mstack_f_.ptr, mult_.ptr, local_.ptr, nhl_.ptr, l0, n, p, kp);
		MLENS_HIP_CHECK(hipEventRecord(e1, stream_));
	}
	else
	{
		if ((long long)mstack_d_.count < nms) mstack_d_.alloc(size_t(nms));
		k_mstack<double><<<unsigned((nms + 255) / 256), 256, 0, stream_>>>(mult_.ptr, l0, n, p, kp, mstack_d_.ptr);
		MLENS_HIP_CHECK(hipEventRecord(e0, stream_));
		k_m2l_mfma<double><<<blocks, 64 * M2L_WAVES, size_t(kp) * kp * sizeof(double), stream_>>>(m2l_ad_.ptr, m2l_src_.ptr,
			mstack_d_.ptr, mult_.ptr, local_.ptr, nhl_.ptr, l0, n, p, kp);
		MLENS_HIP_CHECK(hipEventRecord(e1, stream_));
	}
	MLENS_HIP_CHECK(hipGetLastError());
}

}  // namespace mlens
