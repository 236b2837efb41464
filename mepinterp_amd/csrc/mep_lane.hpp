// mep_lane.hpp -- cross-lane sums and broadcasts on the VALU: DPP row
// patterns, the gfx950 permlane swaps and readlane (no LDS round trip).
// Every function here reads other lanes, so the lanes it names must all be
// active: call them from wave-uniform control flow only.
#pragma once
#include "mep_common.hpp"

__device__ inline double join_d(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// v of the lane that DPP pattern CTRL selects
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
  return join_d(lo, hi);
}
template <int CTRL>
__device__ inline float dpp_d(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
// lanes with bit 5 clear: a(l) + a(l ^ 32); lanes with bit 5 set: b(l) + b(l ^ 32)
__device__ inline double swap32_sum(double a, double b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  return join_d(lo[0], hi[0]) + join_d(lo[1], hi[1]);
}
// the same over lane bit 4 (rows 0, 2 keep a, rows 1, 3 keep b)
__device__ inline double swap16_sum(double a, double b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  return join_d(lo[0], hi[0]) + join_d(lo[1], hi[1]);
}
__device__ inline float swap16_sum(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap((unsigned)__float_as_int(a),
                                                  (unsigned)__float_as_int(b), false, false);
  return __int_as_float((int)r[0]) + __int_as_float((int)r[1]);
}
// the same over lane bit 3 (hi = bit 3 set)
__device__ inline double ror8_sum(double a, double b, bool hi) {
  const double send = hi ? a : b;
  return (hi ? b : a) + dpp_d<0x128>(send);  // row_ror:8
}
// sum over the lanes of each aligned group of 8 (every lane receives it)
template <class T>
__device__ inline T group8_sum(T v) {
  v += dpp_d<0x141>(v);  // row_half_mirror
  v += dpp_d<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_d<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  return v;
}
// sum over the 16 lanes of each row (every lane receives it)
template <class T>
__device__ inline T row16_sum(T v) {
  v += dpp_d<0x128>(v);  // row_ror:8
  v += dpp_d<0x124>(v);  // row_ror:4
  v += dpp_d<0x122>(v);  // row_ror:2
  v += dpp_d<0x121>(v);  // row_ror:1
  return v;
}
// v(l) + v(l ^ 16) + v(l ^ 32) + v(l ^ 48) (every lane receives it)
__device__ inline double rows4_sum(double v) {
  v = swap16_sum(v, v);
  return swap32_sum(v, v);
}
// sum over each aligned team of W lanes (W = 8, 16, 32; every lane receives it)
template <int W, class T>
__device__ inline T team_sum(T v) {
  static_assert(W == 8 || W == 16 || W == 32, "team width");
  if constexpr (W == 8) {
    return group8_sum(v);
  } else {
    v = row16_sum(v);
    if constexpr (W == 32) v = swap16_sum(v, v);
    return v;
  }
}
// v of lane `src` of the caller's team of W lanes (src uniform)
__device__ inline double readlane_d(double v, int lane) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
  return join_d(lo, hi);
}
__device__ inline float readlane_d(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template <int W, class T>
__device__ inline T team_bcast(T v, int src) {
  const int tm = (threadIdx.x & 63) / W;
  T r = readlane_d(v, src);
#pragma unroll
  for (int m = 1; m < 64 / W; ++m) {
    const T o = readlane_d(v, m * W + src);
    r = (tm == m) ? o : r;
  }
  return r;
}
// Sums of 24 values over lane bits 3-5, scattered: lane l ends with the
// slots 12 b5 + 6 b4 + 3 b3 + s (s < 3) of l's bits b5 b4 b3, each summed
// over the 8 lanes that share l's bits 0-2.
__device__ inline void scatter_sum24(const double* v, double* o, int lane) {
  double h[12], q[6];
#pragma unroll
  for (int j = 0; j < 12; ++j) h[j] = swap32_sum(v[j], v[j + 12]);
#pragma unroll
  for (int j = 0; j < 6; ++j) q[j] = swap16_sum(h[j], h[j + 6]);
  const bool b3 = (lane & 8) != 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) o[j] = ror8_sum(q[j], q[j + 3], b3);
}
