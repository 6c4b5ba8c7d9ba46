// Wave-level building blocks shared by the block-tridiagonal kernels: DPP reductions and lane broadcasts
// through scalar registers (no LDS round trips).
#pragma once
#include <hip/hip_runtime.h>

namespace gfa {
namespace wave {

// Maximum of a 32-bit unsigned value over the wave by DPP row shifts and row broadcasts (no LDS round trips).
__device__ __forceinline__ unsigned dpp_wave_max_u32(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false));  // row_shr:1
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false));  // row_shr:2
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false));  // row_shr:4
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false));  // row_shr:8
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast:15
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Sum of a double over the 64 lanes, returned to every lane: an inclusive scan by DPP row shifts and row
// broadcasts, read from lane 63. The summation order is fixed, so results are reproducible.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_add_f64(double v) {
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
  return v + __hiloint2double(hi, lo);
}
__device__ __forceinline__ double dpp_wave_sum(double v) {
  v = dpp_add_f64<0x111, 0xF>(v);  // row_shr:1
  v = dpp_add_f64<0x112, 0xF>(v);  // row_shr:2
  v = dpp_add_f64<0x114, 0xF>(v);  // row_shr:4
  v = dpp_add_f64<0x118, 0xF>(v);  // row_shr:8
  v = dpp_add_f64<0x142, 0xA>(v);  // row_bcast:15
  v = dpp_add_f64<0x143, 0xC>(v);  // row_bcast:31
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// First lane holding the largest a >= 0 among the candidate lanes (LAPACK idamax semantics); -1 if none.
// Non-negative doubles order like their bit patterns, so the maximum is found on the high word, then on the
// low word among the lanes that hold the maximal high word (skipped when only one lane holds it).
__device__ __forceinline__ int wave_first_argmax(double a, bool cand) {
  const unsigned long long bits = cand ? (unsigned long long)__double_as_longlong(a) : 0ull;
  const unsigned hi = (unsigned)(bits >> 32), lo = (unsigned)bits;
  const unsigned mh = dpp_wave_max_u32(hi);
  const unsigned long long top = __ballot(cand && hi == mh);
  if (__popcll(top) == 1) return __ffsll((long long)top) - 1;  // the usual case: one lane holds the top word
  const unsigned ml = dpp_wave_max_u32(hi == mh ? lo : 0u);
  const unsigned long long hit = __ballot(cand && hi == mh && lo == ml);
  return hit ? __ffsll((long long)hit) - 1 : -1;
}

// The value of x held by lane l (l uniform in the wave), through scalar registers rather than LDS.
__device__ __forceinline__ double lane_value(double x, int l) {
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// First lane holding the largest float a >= 0 among the candidate lanes; -1 if none. One pass: non-negative
// floats order like their bit patterns.
__device__ __forceinline__ int wave_first_argmax(float a, bool cand) {
  const unsigned bits = cand ? (unsigned)__float_as_int(a) : 0u;
  const unsigned mb = dpp_wave_max_u32(bits);
  const unsigned long long hit = __ballot(cand && bits == mb);
  return hit ? __ffsll((long long)hit) - 1 : -1;
}

}  // namespace wave
}  // namespace gfa
