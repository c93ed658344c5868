// Sums across the lanes of a DPP row (16 lanes) without LDS round trips.
#pragma once

#include <hip/hip_runtime.h>

namespace isvd {

// move within a DPP row; 64-bit values move as two 32-bit halves
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the 16 lanes of a row, every lane receiving the same (bitwise) total: pairs
// (quad_perm 1,0,3,2), quads (quad_perm 2,3,0,1), half rows (row_half_mirror), rows
// (row_mirror).  Each step adds two partial sums in both orders, which FP addition makes equal.
// All 16 lanes of the row must be active.
template <typename T>
__device__ __forceinline__ T row16_sum(T v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return v;
}

// Sum over aligned groups of `width` lanes (a power of two <= 64; uniform across the wave),
// every lane receiving the total.
template <typename T>
__device__ __forceinline__ T lane_group_sum(T v, int width) {
  if (width >= 2) v += dpp_mov<0xB1>(v);
  if (width >= 4) v += dpp_mov<0x4E>(v);
  if (width >= 8) v += dpp_mov<0x141>(v);
  if (width >= 16) v += dpp_mov<0x140>(v);
  if (width >= 32) v += __shfl_xor(v, 16, width);
  if (width >= 64) v += __shfl_xor(v, 32, width);
  return v;
}

}  // namespace isvd
