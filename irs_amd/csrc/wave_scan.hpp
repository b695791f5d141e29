// Inclusive scans over a 64-lane wave with DPP row shifts and row broadcasts:
// a few cycles per step, where a lane shuffle (ds_bpermute) waits on LDS.
#pragma once

#include <hip/hip_runtime.h>

namespace mlens
{

/*
Inclusive scans over the wave: within each row of 16 lanes by shifts of 1,
2, 4 and 8, then across rows by broadcasting lane 15 into rows 1 and 3 and
lane 31 into rows 2 and 3. A lane whose source lies outside its row, or whose
row is masked off, takes the identity.
*/
__device__ __forceinline__ int wave_prefix_sum(int v)
{
	v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);   // row_shr:1
	v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);   // row_shr:2
	v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);   // row_shr:4
	v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);   // row_shr:8
	v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15
	v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31
	return v;
}

// one DPP step on an FP64 value (both halves moved alike; identity 0.0)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v)
{
	const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
	const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b)), CTRL, ROW_MASK, 0xf, false));
	const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), CTRL, ROW_MASK, 0xf, false));
	return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// the same inclusive sum in FP64
__device__ __forceinline__ double wave_prefix_sum(double v)
{
	v += dpp_f64<0x111, 0xf>(v);   // row_shr:1
	v += dpp_f64<0x112, 0xf>(v);   // row_shr:2
	v += dpp_f64<0x114, 0xf>(v);   // row_shr:4
	v += dpp_f64<0x118, 0xf>(v);   // row_shr:8
	v += dpp_f64<0x142, 0xa>(v);   // row_bcast:15
	v += dpp_f64<0x143, 0xc>(v);   // row_bcast:31
	return v;
}

// the same for the maximum of values >= -1
__device__ __forceinline__ int wave_max_scan(int v)
{
	v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xf, 0xf, false));   // row_shr:1
	v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xf, 0xf, false));   // row_shr:2
	v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xf, 0xf, false));   // row_shr:4
	v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xf, 0xf, false));   // row_shr:8
	v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xa, 0xf, false));   // row_bcast:15
	v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xc, 0xf, false));   // row_bcast:31
	return v;
}

// maximum over aligned groups of g lanes (g a power of two, the same in every
// lane): quad permutations, then half-row and row mirrors, then shuffles
__device__ __forceinline__ int wave_group_max(int m, int g)
{
	const int lowest = -2147483647 - 1;
	if (g > 1) m = max(m, __builtin_amdgcn_update_dpp(lowest, m, 0xB1, 0xf, 0xf, false));    // quad_perm [1,0,3,2]
	if (g > 2) m = max(m, __builtin_amdgcn_update_dpp(lowest, m, 0x4E, 0xf, 0xf, false));    // quad_perm [2,3,0,1]
	if (g > 4) m = max(m, __builtin_amdgcn_update_dpp(lowest, m, 0x141, 0xf, 0xf, false));   // row_half_mirror
	if (g > 8) m = max(m, __builtin_amdgcn_update_dpp(lowest, m, 0x140, 0xf, 0xf, false));   // row_mirror
	if (g > 16) m = max(m, __shfl_xor(m, 16));
	if (g > 32) m = max(m, __shfl_xor(m, 32));
	return m;
}

}  // namespace mlens
